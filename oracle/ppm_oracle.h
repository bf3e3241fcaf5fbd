/*
 * ppm_oracle.h -- CPU restatement of PPM's particle<->mesh interpolation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in ppm_legacy_amd/: only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  Nothing in the product
 * path links, calls or falls back to it.
 *
 * Every routine follows the reference Fortran literally (same formulas, same
 * evaluation order, same index conventions, FP64 or FP32 as the reference's
 * double/single specifics); each carries the reference file:line it restates.
 * Paths are relative to the reference tree's ppm/src/.
 *
 * Conventions shared with the reference:
 *   - positions xp(dim,Np): AoS, coordinate fastest (Fortran column major);
 *   - strengths up(lda,Np): AoS, component fastest;
 *   - field_up(lda, 1-g:Nm+g, 1-g:Nm+g [, 1-g:Nm+g]): component fastest, then
 *     x, y, z; Fortran node index i (1-based, lower bound 1-g) sits at
 *     min_phys + (i-1)*h.  One sub covering the whole domain (istart = 1,
 *     nnodes = Nm), as PPM builds on a single process.
 *   - info: 0 on success, else the ppm_err_* code (ppm_error.h numbering).
 */
#ifndef PPM_ORACLE_H
#define PPM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_topo {
    int    dim;          /* 2 or 3 */
    int    Nm[3];        /* mesh nodes per axis (both boundary nodes counted) */
    double min_phys[3];
    double max_phys[3];
    int    bcdef[6];     /* per face x-,x+,y-,y+,z-,z+ ; 1 = periodic (ppm_param.h:333) */
} or_topo;

#define OR_OK              0
#define OR_ERR_PART_UNASS  405   /* ppm_err_part_unass, ppm_error.h:217 */
#define OR_ERR_ARGUMENT    503   /* ppm_err_argument,   ppm_error.h:237 */
#define OR_ERR_ALLOC       101   /* ppm_err_alloc,      ppm_error.h:161 */
#define OR_ERR_BOUNDS      9001  /* oracle only: stencil node outside the padded field */

/* element kinds of or_ghost_kind (the reference's _d, _s, _i, _l, _dc, _sc
 * specifics of ppm_map_field_ghost); logicals are 32-bit, nonzero = .TRUE. */
#define OR_KIND_D  0
#define OR_KIND_S  2
#define OR_KIND_I  3
#define OR_KIND_L  4
#define OR_KIND_DC 5
#define OR_KIND_SC 6

#define OR_KERNEL_BSP2 1         /* ppm_param_rmsh_kernel_bsp2, ppm_param.h:405 */
#define OR_KERNEL_MP4  2         /* ppm_param_rmsh_kernel_mp4,  ppm_param.h:406 */

/* Double (d) and single (s) specifics, mirroring the reference's
 * ppm_interp_p2m_{d,s}{s,v}_{2d,3d} family.  lda = 1 covers the scalar forms. */
#define OR_DECLARE(SFX, REAL)                                                          \
int or_p2m##SFX(const or_topo* t, int kernel, int np, const REAL* xp, int lda,         \
                const REAL* up, const int* ghost, REAL* field, int fold);              \
int or_m2p##SFX(const or_topo* t, int kernel, int np, const REAL* xp, int lda,         \
                const REAL* field, const int* ghost, REAL* up, int ghost_fill);        \
int or_ghost_put##SFX(const or_topo* t, int lda, const int* ghost, REAL* field);       \
int or_ghost_get##SFX(const or_topo* t, int lda, const int* ghost, REAL* field);       \
int or_rank3d##SFX(int np, const REAL* xp, const REAL* xmin, const REAL* xmax,         \
                   const int* Nm, const int* Ngl, int* lpdx, int* lhbx, int* icount,   \
                   int* nout, int* ncorr);                                             \
int or_rank2d##SFX(int np, const REAL* xp, const REAL* xmin, const REAL* xmax,         \
                   const int* Nm, const int* Ngl, int* lpdx, int* lhbx, int* icount,   \
                   int* nout, int* ncorr);                                             \
int or_impose_part_bc##SFX(const or_topo* t, int np, REAL* xp);                        \
int or_comp_weights##SFX(const or_topo* t, int kernel, int np, const REAL* xp,         \
                         REAL* wx1, REAL* wx2, REAL* wx3);                             \
int or_remesh##SFX(const or_topo* t, int kernel, int np, const REAL* xp, int lda,      \
                   const REAL* up, const int* ghost, const REAL* wx1, const REAL* wx2, \
                   const REAL* wx3, REAL* field);                                      \
int or_create_part##SFX(const or_topo* t, int lda, const REAL* field, const int* ghost,\
                        const REAL* cutoff, const REAL* cutoff_weights, int resetpos,  \
                        int maxnp, int* np_out, REAL* xp, REAL* up);                   \
int or_p2m_adjoint##SFX(const or_topo* t, int kernel, int np, const REAL* xp, int lda, \
                        const REAL* up, const REAL* g, REAL* grad_u, REAL* grad_x);

OR_DECLARE(_d, double)
OR_DECLARE(_s, float)

/* ppm_map_field_ghost with ghost_get (get = 1) or ghost_put (get = 0), push,
 * send and pop, for element kind OR_KIND_*, on the padded field
 * (lda, 1-g:Nm+g, ...); mask: NULL or one logical per stored node. */
int or_ghost_kind(const or_topo* t, int kind, int lda, const int* ghost, void* field, const int* mask, int get);

/* Multi-threaded FP64 P2M (slab decomposition in z over OpenMP threads, each
 * thread scattering its particles into a private slab with ghost layers, then
 * the ghost add-exchange of ppm_map_field_ghost_put and the periodic fold).
 * It is the CPU baseline's analogue of running the reference with one MPI
 * rank per core. */
int or_p2m_par_d(const or_topo* t, int kernel, int np, const double* xp, int lda,
                 const double* up, const int* ghost, double* field, int nthreads);
int or_m2p_par_d(const or_topo* t, int kernel, int np, const double* xp, int lda,
                 const double* field, const int* ghost, double* up, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
