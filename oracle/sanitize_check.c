/*
 * sanitize_check.c -- runs every routine of the CPU oracle on small inputs
 * under AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle
 * sanitize).  TEST INFRASTRUCTURE ONLY, like the oracle itself.
 *
 * Cases: 2D and 3D, both kernels, FP64 and FP32, periodic and open axes,
 * ghost widths 1..3, empty input.  Every call must return 0 and, where the
 * result has a closed form, match it: P2M conserves the strengths on the
 * real nodes of a periodic mesh, M2P of a constant field returns the
 * constant.  Exits non-zero on the first failure; the sanitizers abort on the
 * first invalid access or undefined operation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ppm_oracle.h"

static int failures = 0;

#define CHECK(cond, ...)                     \
    do {                                     \
        if (!(cond)) {                       \
            fprintf(stderr, __VA_ARGS__);    \
            fputc('\n', stderr);             \
            ++failures;                      \
        }                                    \
    } while (0)

static unsigned long long rng_state = 88172645463325252ull;
static double uniform(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static long field_len(const or_topo* t, int lda, const int* g) {
    long n = lda;
    for (int d = 0; d < t->dim; ++d) n *= t->Nm[d] + 2 * g[d];
    return n;
}

/* Sum of component 0 over the real unique nodes 1..Nm-1 (periodic axes). */
static double real_sum(const or_topo* t, int lda, const int* g, const double* f) {
    const int e0 = t->Nm[0] + 2 * g[0], e1 = t->Nm[1] + 2 * g[1];
    const int nz = t->dim == 3 ? t->Nm[2] - 1 : 1;
    double s = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < t->Nm[1] - 1; ++j)
            for (int i = 0; i < t->Nm[0] - 1; ++i) {
                const long z = t->dim == 3 ? k + g[2] : 0;
                s += f[(((z * e1) + (j + g[1])) * (long)e0 + (i + g[0])) * lda];
            }
    return s;
}

static void run_double(int dim, int kernel, int periodic, int gw, int np) {
    or_topo t = {0};
    t.dim = dim;
    for (int d = 0; d < 3; ++d) {
        t.Nm[d] = d < dim ? 9 + d : 2;
        t.min_phys[d] = 0.0;
        t.max_phys[d] = 1.0;
    }
    for (int f = 0; f < 6; ++f) t.bcdef[f] = periodic ? 1 : 2;
    int g[3] = {gw, gw, gw};
    const int lda = 2;
    double* xp = malloc(sizeof(double) * (size_t)(np > 0 ? np : 1) * dim);
    double* up = malloc(sizeof(double) * (size_t)(np > 0 ? np : 1) * lda);
    for (int p = 0; p < np; ++p) {
        for (int d = 0; d < dim; ++d) xp[p * dim + d] = 0.05 + 0.9 * uniform();
        for (int c = 0; c < lda; ++c) up[p * lda + c] = uniform() - 0.5;
    }
    const long nf = field_len(&t, lda, g);
    double* field = calloc((size_t)nf, sizeof(double));
    int info = or_p2m_d(&t, kernel, np, xp, lda, up, g, field, 1);
    CHECK(info == 0, "p2m_d dim %d kernel %d periodic %d g %d: info %d", dim, kernel, periodic, gw, info);
    if (periodic && info == 0) {
        double tot = 0.0;
        for (int p = 0; p < np; ++p) tot += up[p * lda];
        CHECK(fabs(real_sum(&t, lda, g, field) - tot) < 1e-9 * (1.0 + fabs(tot)), "p2m_d mass not conserved");
    }
    double* back = calloc((size_t)(np > 0 ? np : 1) * lda, sizeof(double));
    info = or_ghost_get_d(&t, lda, g, field);
    CHECK(info == 0, "ghost_get_d: info %d", info);
    info = or_ghost_put_d(&t, lda, g, field);
    CHECK(info == 0, "ghost_put_d: info %d", info);
    {
        /* every element kind, with and without a mask (16 bytes per element
         * covers the complex double kind) */
        const long nn = nf / lda;
        void* fk = calloc((size_t)nf, 16);
        int* mk = malloc(sizeof(int) * (size_t)nn);
        for (long i = 0; i < nn; ++i) mk[i] = (int)(i % 3 != 0);
        const int kinds[6] = {OR_KIND_D, OR_KIND_S, OR_KIND_I, OR_KIND_L, OR_KIND_DC, OR_KIND_SC};
        for (int q = 0; q < 6; ++q)
            for (int use_mask = 0; use_mask < 2; ++use_mask)
                for (int get = 0; get < 2; ++get) {
                    info = or_ghost_kind(&t, kinds[q], lda, g, fk, use_mask ? mk : NULL, get);
                    CHECK(info == 0, "ghost_kind %d mask %d get %d: info %d", kinds[q], use_mask, get, info);
                }
        free(fk);
        free(mk);
    }
    for (long i = 0; i < nf; ++i) field[i] = 1.5;
    info = or_m2p_d(&t, kernel, np, xp, lda, field, g, back, 1);
    CHECK(info == 0, "m2p_d: info %d", info);
    for (int p = 0; p < np && info == 0; ++p)
        CHECK(fabs(back[p * lda] - 1.5) < 1e-12, "m2p_d of a constant: %g", back[p * lda]);
    /* comp_weights + remesh */
    const int ks = kernel == OR_KERNEL_MP4 ? 4 : 2;
    double* w[3];
    for (int d = 0; d < 3; ++d) w[d] = calloc((size_t)(np > 0 ? np : 1) * ks, sizeof(double));
    info = or_comp_weights_d(&t, kernel, np, xp, w[0], w[1], w[2]);
    CHECK(info == 0, "comp_weights_d: info %d", info);
    info = or_remesh_d(&t, kernel, np, xp, lda, up, g, w[0], w[1], w[2], field);
    CHECK(info == 0, "remesh_d: info %d", info);
    /* create_part from the remeshed field */
    const double cut[2] = {0.0, 1e300};
    const int maxnp = (int)nf;
    double* xo = calloc((size_t)maxnp * dim, sizeof(double));
    double* uo = calloc((size_t)maxnp * lda, sizeof(double));
    int nout = 0;
    info = or_create_part_d(&t, lda, field, g, cut, NULL, 1, maxnp, &nout, xo, uo);
    CHECK(info == 0, "create_part_d: info %d", info);
    /* adjoint: periodic meshes, cotangent on the unique nodes (the padded
       buffer is larger than that compact array) */
    double* gu = calloc((size_t)(np > 0 ? np : 1) * lda, sizeof(double));
    double* gx = calloc((size_t)(np > 0 ? np : 1) * dim, sizeof(double));
    for (long i = 0; i < nf; ++i) field[i] = uniform();
    if (periodic) {
        info = or_p2m_adjoint_d(&t, kernel, np, xp, lda, up, field, gu, gx);
        CHECK(info == 0, "p2m_adjoint_d: info %d", info);
    }
    /* impose_part_bc on shifted copies */
    for (int p = 0; p < np * dim; ++p) xp[p] += 0.5;
    info = or_impose_part_bc_d(&t, np, xp);
    CHECK(info == 0, "impose_part_bc_d: info %d", info);
    /* rank */
    const double xmin[3] = {0.0, 0.0, 0.0}, xmax[3] = {1.0, 1.0, 1.0};
    const int Nm[3] = {4, 5, 6};
    const int Ngl[6] = {1, 1, 0, 1, 1, 0};
    const int nbox = (4 + 2) * (5 + 1) * (dim == 3 ? 6 + 1 : 1);
    int* lpdx = calloc((size_t)(np > 0 ? np : 1), sizeof(int));
    int* lhbx = calloc((size_t)nbox + 1, sizeof(int));
    int icount = 0, nout2 = 0, ncorr = 0;
    if (dim == 3) info = or_rank3d_d(np, xp, xmin, xmax, Nm, Ngl, lpdx, lhbx, &icount, &nout2, &ncorr);
    else info = or_rank2d_d(np, xp, xmin, xmax, Nm, Ngl, lpdx, lhbx, &icount, &nout2, &ncorr);
    CHECK(info == 0 || info == OR_ERR_PART_UNASS, "rank_d: info %d", info);
    free(lpdx);
    free(lhbx);
    free(gu);
    free(gx);
    free(xo);
    free(uo);
    for (int d = 0; d < 3; ++d) free(w[d]);
    free(back);
    free(field);
    free(xp);
    free(up);
}

static void run_single(int dim, int kernel, int gw, int np) {
    or_topo t = {0};
    t.dim = dim;
    for (int d = 0; d < 3; ++d) {
        t.Nm[d] = d < dim ? 8 : 2;
        t.max_phys[d] = 1.0;
    }
    for (int f = 0; f < 6; ++f) t.bcdef[f] = 1;
    int g[3] = {gw, gw, gw};
    float* xp = malloc(sizeof(float) * (size_t)(np > 0 ? np : 1) * dim);
    float* up = malloc(sizeof(float) * (size_t)(np > 0 ? np : 1));
    for (int p = 0; p < np; ++p) {
        for (int d = 0; d < dim; ++d) xp[p * dim + d] = (float)(0.05 + 0.9 * uniform());
        up[p] = (float)(uniform() - 0.5);
    }
    const long nf = field_len(&t, 1, g);
    float* field = calloc((size_t)nf, sizeof(float));
    int info = or_p2m_s(&t, kernel, np, xp, 1, up, g, field, 1);
    CHECK(info == 0, "p2m_s: info %d", info);
    float* back = calloc((size_t)(np > 0 ? np : 1), sizeof(float));
    info = or_m2p_s(&t, kernel, np, xp, 1, field, g, back, 1);
    CHECK(info == 0, "m2p_s: info %d", info);
    free(back);
    free(field);
    free(xp);
    free(up);
}

int main(void) {
    for (int dim = 2; dim <= 3; ++dim)
        for (int kernel = OR_KERNEL_BSP2; kernel <= OR_KERNEL_MP4; ++kernel)
            for (int periodic = 0; periodic <= 1; ++periodic)
                for (int gw = 2; gw <= 3; ++gw) {
                    run_double(dim, kernel, periodic, gw, 300);
                    run_double(dim, kernel, periodic, gw, 0);
                }
    for (int dim = 2; dim <= 3; ++dim)
        for (int kernel = OR_KERNEL_BSP2; kernel <= OR_KERNEL_MP4; ++kernel) run_single(dim, kernel, 2, 200);
    /* the OpenMP baseline variants */
    or_topo t = {3, {17, 17, 17}, {0, 0, 0}, {1, 1, 1}, {1, 1, 1, 1, 1, 1}};
    const int np = 2000, g[3] = {1, 1, 1};
    double* xp = malloc(sizeof(double) * np * 3);
    double* up = malloc(sizeof(double) * np * 3);
    for (int i = 0; i < np * 3; ++i) {
        xp[i] = uniform();
        up[i] = uniform();
    }
    double* field = calloc((size_t)field_len(&t, 3, g), sizeof(double));
    double* back = calloc((size_t)np * 3, sizeof(double));
    CHECK(or_p2m_par_d(&t, OR_KERNEL_MP4, np, xp, 3, up, g, field, 2) == 0, "p2m_par_d failed");
    CHECK(or_m2p_par_d(&t, OR_KERNEL_MP4, np, xp, 3, field, g, back, 2) == 0, "m2p_par_d failed");
    free(xp);
    free(up);
    free(field);
    free(back);
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("sanitized oracle run: all checks passed\n");
    return 0;
}
