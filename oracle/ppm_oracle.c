/*
 * ppm_oracle.c -- CPU restatement of PPM's particle<->mesh interpolation path.
 * TEST INFRASTRUCTURE ONLY (see ppm_oracle.h): the checker for the HIP path,
 * never linked into it.
 *
 * Build: oracle/Makefile (gcc -O2, no -march so no FMA contraction: the
 * floating-point evaluation order is the reference's).
 */
#include "ppm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Geometry of a padded field field_up(lda, 1-g:Nm+g, ...), the reference's
 * allocation bounds (ppm_rmsh_remesh.f:523-571). */
typedef struct { int dim, lda; int lo[3]; int ext[3]; } or_fgeom;

static or_fgeom or_fgeom_make(const or_topo* t, int lda, const int* g)
{
    or_fgeom f;
    f.dim = t->dim;
    f.lda = lda;
    for (int d = 0; d < 3; ++d) {
        if (d < t->dim) { f.lo[d] = 1 - g[d]; f.ext[d] = t->Nm[d] + 2 * g[d]; }
        else { f.lo[d] = 1; f.ext[d] = 1; }
    }
    return f;
}

static int64_t or_fgeom_size(const or_fgeom* f)
{
    return (int64_t)f->lda * f->ext[0] * f->ext[1] * f->ext[2];
}

/* Offset of Fortran node (i,j,k); returns 1 when it lies outside the field. */
static int or_fidx(const or_fgeom* f, int i, int j, int k, int64_t* off)
{
    int a = i - f->lo[0], b = j - f->lo[1], c = f->dim == 3 ? k - f->lo[2] : 0;
    if (a < 0 || a >= f->ext[0] || b < 0 || b >= f->ext[1] || c < 0 || c >= f->ext[2]) return 1;
    *off = (int64_t)f->lda * (a + (int64_t)f->ext[0] * (b + (int64_t)f->ext[1] * c));
    return 0;
}

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

#define REAL double
#define SFX _d
#define TRUNC trunc
#define FABS fabs
#define FLOOR floor
#define SQRT sqrt
#include "ppm_oracle_body.inc"
#undef REAL
#undef SFX
#undef TRUNC
#undef FABS
#undef FLOOR
#undef SQRT

#define REAL float
#define SFX _s
#define TRUNC truncf
#define FABS fabsf
#define FLOOR floorf
#define SQRT sqrtf
#include "ppm_oracle_body.inc"
#undef REAL
#undef SFX
#undef TRUNC
#undef FABS
#undef FLOOR
#undef SQRT

/* ---------------------------------------------------------------------------
 * ppm_map_field_ghost (ghost_get / ghost_put, push, send, pop on one sub) for
 * every element kind of the reference's specifics, with the optional mask.
 * Push packs every node of every block into the double send buffer
 * (ppm_map_field_push_3d.f:556-592: REAL(.,ppm_kind_double) of reals and
 * integers, real and imaginary part of complex values, 1.0 / 0.0 of
 * logicals; its mask test is commented out at :546-549); the pop reads the
 * buffer back block by block in list order (ppm_map_field_pop_3d.f:
 * 1818-1897, pop_2d.f:1757-1834).  Where the mask (one logical per stored
 * node of the padded box) is .FALSE. the node keeps its value and the buffer
 * steps past it.  Replace converts back (INT() of integers, value >
 * 1 - ppm_myepsd of logicals); add adds (integers, reals, complex parts) or
 * takes .AND. (logicals).  ghost_get packs the real block (start) and
 * replaces its images (start + off); ghost_put packs the images and adds them
 * onto the real block (ppm_map_field_ghost_put.f:133-134,242-281).
 * ------------------------------------------------------------------------- */
int or_ghost_kind(const or_topo* t, int kind, int lda, const int* g, void* field, const int* mask, int get)
{
    /* ppm_myepsd = 10**tolexp (ppm_init.f:356); any value in (0,1) reads the
     * packed 1.0 / 0.0 the same way */
    const double myeps = 1e-15;
    int parts;
    switch (kind) {
    case OR_KIND_D: case OR_KIND_S: case OR_KIND_I: case OR_KIND_L: parts = 1; break;
    case OR_KIND_DC: case OR_KIND_SC: parts = 2; break;
    default: return OR_ERR_ARGUMENT;
    }
    blk_d b[26];
    int nb = ghost_blocks_d(t, g, b);
    or_fgeom f = or_fgeom_make(t, lda, g);
    int64_t total = 0;
    for (int q = 0; q < nb; ++q) total += (int64_t)b[q].size[0] * b[q].size[1] * b[q].size[2] * lda * parts;
    double* buf = (double*)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
    if (!buf) return OR_ERR_ALLOC;
    double* fd = (double*)field;
    float* fs = (float*)field;
    int* fi = (int*)field;
    int64_t p = 0;
    for (int q = 0; q < nb; ++q) {
        const int sh = get ? 0 : 1;   /* put packs the image block */
        for (int k = 0; k < b[q].size[2]; ++k)
            for (int j = 0; j < b[q].size[1]; ++j)
                for (int i = 0; i < b[q].size[0]; ++i) {
                    int64_t off;
                    or_fidx(&f, b[q].start[0] + sh * b[q].off[0] + i, b[q].start[1] + sh * b[q].off[1] + j,
                            b[q].start[2] + sh * b[q].off[2] + k, &off);
                    for (int c = 0; c < lda; ++c) {
                        const int64_t e = off + c;
                        switch (kind) {
                        case OR_KIND_D: buf[p++] = fd[e]; break;
                        case OR_KIND_S: buf[p++] = (double)fs[e]; break;
                        case OR_KIND_I: buf[p++] = (double)fi[e]; break;
                        case OR_KIND_L: buf[p++] = fi[e] ? 1.0 : 0.0; break;
                        case OR_KIND_DC: buf[p++] = fd[2 * e]; buf[p++] = fd[2 * e + 1]; break;
                        default: buf[p++] = (double)fs[2 * e]; buf[p++] = (double)fs[2 * e + 1]; break;
                        }
                    }
                }
    }
    p = 0;
    for (int q = 0; q < nb; ++q) {
        const int sh = get ? 1 : 0;   /* get writes the image block */
        for (int k = 0; k < b[q].size[2]; ++k)
            for (int j = 0; j < b[q].size[1]; ++j)
                for (int i = 0; i < b[q].size[0]; ++i) {
                    int64_t off;
                    or_fidx(&f, b[q].start[0] + sh * b[q].off[0] + i, b[q].start[1] + sh * b[q].off[1] + j,
                            b[q].start[2] + sh * b[q].off[2] + k, &off);
                    if (mask && !mask[off / lda]) {
                        p += (int64_t)lda * parts;   /* skip this point */
                        continue;
                    }
                    for (int c = 0; c < lda; ++c) {
                        const int64_t e = off + c;
                        switch (kind) {
                        case OR_KIND_D: fd[e] = get ? buf[p] : fd[e] + buf[p]; p++; break;
                        case OR_KIND_S: fs[e] = get ? (float)buf[p] : fs[e] + (float)buf[p]; p++; break;
                        case OR_KIND_I: {
                            const int v = (int)buf[p++];   /* INT(): exact for every 32-bit value */
                            fi[e] = get ? v : (int)((unsigned)fi[e] + (unsigned)v);
                            break;
                        }
                        case OR_KIND_L: {
                            const int v = buf[p++] > 1.0 - myeps;
                            fi[e] = get ? v : (fi[e] != 0 && v);
                            break;
                        }
                        case OR_KIND_DC:
                            fd[2 * e] = get ? buf[p] : fd[2 * e] + buf[p];
                            fd[2 * e + 1] = get ? buf[p + 1] : fd[2 * e + 1] + buf[p + 1];
                            p += 2;
                            break;
                        default:
                            fs[2 * e] = get ? (float)buf[p] : fs[2 * e] + (float)buf[p];
                            fs[2 * e + 1] = get ? (float)buf[p + 1] : fs[2 * e + 1] + (float)buf[p + 1];
                            p += 2;
                            break;
                        }
                    }
                }
    }
    free(buf);
    return OR_OK;
}

/* ---------------------------------------------------------------------------
 * Multi-threaded FP64 P2M for the CPU baseline.  The reference parallelises
 * P2M by spatial decomposition: each MPI rank scatters the particles of its
 * subs into its own field with ghost layers (ppm_interp_p2m.f:783-3567), then
 * ppm_map_field_ghost_put adds the ghost layers onto the neighbours' nodes
 * (ppm_map_field_send.f:446-545, ppm_map_field_pop_3d.f:781-800).  Here each
 * thread owns a z-slab of cells, scatters into a private slab buffer with
 * ghost planes, and the add-exchange + periodic fold then runs plane by plane
 * (each output plane summed by one thread).  Fully periodic 3D meshes only.
 * ------------------------------------------------------------------------- */
int or_p2m_par_d(const or_topo* t, int kernel, int np, const double* xp, int lda,
                 const double* up, const int* g, double* field, int nthreads)
{
    if (t->dim != 3 || lda < 1 || lda > 16) return OR_ERR_ARGUMENT;
    for (int d = 0; d < 6; ++d) if (t->bcdef[d] != 1) return OR_ERR_ARGUMENT;
    if (kernel != OR_KERNEL_MP4 && kernel != OR_KERNEL_BSP2) return OR_ERR_ARGUMENT;
    for (int p = 0; p < np; ++p)
        if (!in_sub_d(t, xp + (size_t)p * 3)) return OR_ERR_PART_UNASS;
    if (nthreads < 1) nthreads = 1;
    int N[3] = { t->Nm[0] - 1, t->Nm[1] - 1, t->Nm[2] - 1 };
    if (nthreads > N[2]) nthreads = N[2];
    double dxi[3];
    spacing_mk_d(t, dxi);
    /* slab of cells [zlo[s], zlo[s+1]) per thread; particle buckets by slab */
    int* zlo = (int*)malloc(sizeof(int) * (size_t)(nthreads + 1));
    for (int s = 0; s <= nthreads; ++s) zlo[s] = (int)((int64_t)N[2] * s / nthreads);
    int* slab_of = (int*)malloc(sizeof(int) * (size_t)(N[2] + 1));
    for (int s = 0; s < nthreads; ++s)
        for (int z = zlo[s]; z < zlo[s + 1]; ++z) slab_of[z] = s;
    slab_of[N[2]] = nthreads - 1;   /* INT(x0) may reach N through rounding */
    int64_t* cnt = (int64_t*)calloc((size_t)nthreads + 1, sizeof(int64_t));
    int* pslab = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    for (int p = 0; p < np; ++p) {
        int cz = (int)((xp[(size_t)p * 3 + 2] - t->min_phys[2]) * dxi[2]);
        pslab[p] = slab_of[cz];
        cnt[pslab[p] + 1]++;
    }
    for (int s = 0; s < nthreads; ++s) cnt[s + 1] += cnt[s];
    int* order = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)nthreads);
    for (int s = 0; s < nthreads; ++s) cur[s] = cnt[s];
    for (int p = 0; p < np; ++p) order[cur[pslab[p]]++] = p;
    /* private slab buffers: nodes x,y in [-1, N+2], z in [zlo-1, zhi+2] (0-based) */
    int ex = N[0] + 4, ey = N[1] + 4;
    double** buf = (double**)calloc((size_t)nthreads, sizeof(double*));
    int rc = OR_OK;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int ss = 0; ss < nthreads; ++ss) {
        int ez = zlo[ss + 1] - zlo[ss] + 4;
        double* b = (double*)calloc((size_t)lda * ex * ey * ez, sizeof(double));
        buf[ss] = b;
        if (!b) { rc = OR_ERR_ALLOC; continue; }
        for (int64_t q = cnt[ss]; q < cnt[ss + 1]; ++q) {
            int p = order[q];
            const double* x = xp + (size_t)p * 3;
            const double* u = up + (size_t)p * lda;
            double x0[3];
            for (int d = 0; d < 3; ++d) x0[d] = (x[d] - t->min_phys[d]) * dxi[d];
            int lo[3], ks;
            double a[3][4];
            if (kernel == OR_KERNEL_MP4) {
                ks = 4;
                for (int d = 0; d < 3; ++d) {
                    int ip = (int)x0[d];
                    mp4_unrolled_d(x0[d] - (double)ip, a[d]);
                    lo[d] = ip - 1;                 /* 0-based node of a[d][0] */
                }
            } else {
                ks = 2;
                for (int d = 0; d < 3; ++d) {
                    int ip = (int)x0[d] + 1;
                    double x10 = x0[d] - (double)(ip - 1);
                    a[d][0] = 1.0 - x10;
                    a[d][1] = 1.0 + (x10 - 1.0);
                    lo[d] = ip - 1;
                }
            }
            for (int ia = 0; ia < ks; ++ia)
                for (int ib = 0; ib < ks; ++ib)
                    for (int ic = 0; ic < ks; ++ic) {
                        double w = a[0][ia] * a[1][ib] * a[2][ic];
                        int bi = lo[0] + ia + 1, bj = lo[1] + ib + 1;
                        int bk = lo[2] + ic - (zlo[ss] - 1);
                        double* dst = b + (size_t)lda * (bi + (size_t)ex * (bj + (size_t)ey * bk));
                        for (int c = 0; c < lda; ++c) dst[c] = dst[c] + w * u[c];
                    }
        }
    }
    if (rc == OR_OK) {
        /* add-exchange and periodic fold, one output plane per iteration */
        or_fgeom f = or_fgeom_make(t, lda, g);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int k = 0; k < N[2]; ++k) {
            double* plane = (double*)calloc((size_t)lda * N[0] * N[1], sizeof(double));
            for (int s = 0; s < nthreads; ++s) {
                int ez = zlo[s + 1] - zlo[s] + 4;
                for (int bk = 0; bk < ez; ++bk) {
                    int z = bk + zlo[s] - 1;
                    int zz = ((z % N[2]) + N[2]) % N[2];
                    if (zz != k) continue;
                    for (int bj = 0; bj < ey; ++bj) {
                        int yy = (((bj - 1) % N[1]) + N[1]) % N[1];
                        for (int bi = 0; bi < ex; ++bi) {
                            int xx = (((bi - 1) % N[0]) + N[0]) % N[0];
                            const double* src = buf[s] + (size_t)lda * (bi + (size_t)ex * (bj + (size_t)ey * bk));
                            double* dst = plane + (size_t)lda * (xx + (size_t)N[0] * yy);
                            for (int c = 0; c < lda; ++c) dst[c] += src[c];
                        }
                    }
                }
            }
            /* write every field plane that is an image of unique plane k */
            for (int kz = f.lo[2]; kz < f.lo[2] + f.ext[2]; ++kz) {
                int zz = (((kz - 1) % N[2]) + N[2]) % N[2];
                if (zz != k) continue;
                for (int jy = f.lo[1]; jy < f.lo[1] + f.ext[1]; ++jy) {
                    int yy = (((jy - 1) % N[1]) + N[1]) % N[1];
                    for (int ix = f.lo[0]; ix < f.lo[0] + f.ext[0]; ++ix) {
                        int xx = (((ix - 1) % N[0]) + N[0]) % N[0];
                        int64_t off;
                        or_fidx(&f, ix, jy, kz, &off);
                        const double* src = plane + (size_t)lda * (xx + (size_t)N[0] * yy);
                        for (int c = 0; c < lda; ++c) field[off + c] = src[c];
                    }
                }
            }
            free(plane);
        }
    }
    for (int s = 0; s < nthreads; ++s) free(buf[s]);
    free(buf); free(zlo); free(slab_of); free(cnt); free(pslab); free(order); free(cur);
    return rc;
}

/* Multi-threaded FP64 M2P for the CPU baseline: ghost_get on a private copy,
 * then the gather of ppm_interp_m2p.f:895-1240 split over particles. */
int or_m2p_par_d(const or_topo* t, int kernel, int np, const double* xp, int lda,
                 const double* field_in, const int* g, double* up, int nthreads)
{
    if (t->dim != 3 || lda < 1) return OR_ERR_ARGUMENT;
    if (kernel != OR_KERNEL_MP4 && kernel != OR_KERNEL_BSP2) return OR_ERR_ARGUMENT;
    or_fgeom f = or_fgeom_make(t, lda, g);
    int64_t fsz = or_fgeom_size(&f);
    double* field = (double*)malloc(sizeof(double) * (size_t)fsz);
    if (!field) return OR_ERR_ALLOC;
    memcpy(field, field_in, sizeof(double) * (size_t)fsz);
    or_ghost_get_d(t, lda, g, field);
    double dxi[3];
    spacing_single_d(t, dxi);
    int rc = OR_OK;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int p = 0; p < np; ++p) {
        const double* x = xp + (size_t)p * 3;
        double* u = up + (size_t)p * lda;
        for (int c = 0; c < lda; ++c) u[c] = 0.0;
        double x0[3];
        for (int d = 0; d < 3; ++d) x0[d] = (x[d] - t->min_phys[d]) * dxi[d];
        int lo[3], ks;
        double a[3][4];
        if (kernel == OR_KERNEL_MP4) {
            ks = 4;
            for (int d = 0; d < 3; ++d) {
                int ip = (int)x0[d];
                mp4_unrolled_d(x0[d] - (double)ip, a[d]);
                lo[d] = ip;
            }
        } else {
            ks = 2;
            for (int d = 0; d < 3; ++d) {
                int ip = (int)x0[d] + 1;
                double x10 = x0[d] - (double)(ip - 1);
                a[d][0] = 1.0 - x10;
                a[d][1] = 1.0 + (x10 - 1.0);
                lo[d] = ip;
            }
        }
        for (int ia = 0; ia < ks; ++ia)
            for (int ib = 0; ib < ks; ++ib)
                for (int ic = 0; ic < ks; ++ic) {
                    int64_t off;
                    if (or_fidx(&f, lo[0] + ia, lo[1] + ib, lo[2] + ic, &off)) { rc = OR_ERR_BOUNDS; continue; }
                    double w = a[0][ia] * a[1][ib] * a[2][ic];
                    for (int c = 0; c < lda; ++c) u[c] = u[c] + w * field[off + c];
                }
    }
    free(field);
    return rc;
}
