/* ppmg_c_check.c -- a C client of include/ppmg.h, built with gcc against
 * libppmg.so: what a C (or cgo / FFI) caller of the library sees.
 *
 *   ppmg_c_check cpu   error paths that need no device: plan lifecycle,
 *                      registry calls before ppmg_init (903), an unknown
 *                      interpolation kernel (503), ppmg_last_error text
 *   ppmg_c_check gpu   the registry entry points on host arrays:
 *                      ppmg_init, ppmg_mktopo_field, ppmg_interp_p2m and
 *                      ppmg_interp_m2p for both kernels, checked against
 *                      properties the mathematics fixes:
 *                        - the periodic fold conserves mass: the unique
 *                          real nodes sum to the total strength;
 *                        - after the fold the duplicate face node Nm holds
 *                          node 1's value (bitwise);
 *                        - M2P of a constant field returns the constant
 *                          (the weights are a partition of unity).
 *
 * Prints one line per check and "ok" at the end; exits non-zero on the
 * first failure.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ppmg.h"

static int fail(const char* what) {
    int code = 0;
    char msg[512];
    int info = ppmg_last_error(&code, msg, (int)sizeof msg);
    printf("FAIL %s (last error info=%d code=%d: %s)\n", what, info, code, msg);
    return 1;
}

static uint64_t lcg_state = 0x9E3779B97F4A7C15ull;
static double uniform01(void) {
    lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg_state >> 11) * (1.0 / 9007199254740992.0);
}

static int check_cpu(void) {
    ppmg_plan* plan = NULL;
    if (ppmg_plan_create(&plan) != 0 || plan == NULL) return fail("ppmg_plan_create");

    /* a registry call before ppmg_init: ppm_noinit (903) */
    const double mn[3] = {0, 0, 0}, mx[3] = {1, 1, 1};
    const int bc[6] = {1, 1, 1, 1, 1, 1}, nm[3] = {9, 9, 9}, g[3] = {1, 1, 1};
    int topo = 0, mesh = 0, code = 0;
    char msg[512];
    int info = ppmg_mktopo_field(3, mn, mx, bc, nm, g, &topo, &mesh);
    ppmg_last_error(&code, msg, (int)sizeof msg);
    if (info != PPMG_ERROR_ERROR || code != PPMG_ERR_NOINIT) return fail("mktopo before init is 903");
    printf("noinit: info=%d code=%d\n", info, code);

    /* kernel 3 is not an interpolation kernel: ppm_err_argument (503) */
    ppmg_mesh m;
    memset(&m, 0, sizeof m);
    m.dim = 3;
    for (int d = 0; d < 3; ++d) {
        m.nm[d] = 9;
        m.max_phys[d] = 1.0;
    }
    for (int f = 0; f < 6; ++f) m.bcdef[f] = 1;
    ppmg_field fld;
    memset(&fld, 0, sizeof fld);
    fld.lda = 1;
    fld.data = (void*)8;
    info = ppmg_p2m(&m, 3, PPMG_PREC_F64, 1, (const void*)8, 1, (const void*)8, &fld, plan, NULL);
    ppmg_last_error(&code, msg, (int)sizeof msg);
    if (info != PPMG_ERROR_ERROR || code != PPMG_ERR_ARGUMENT || msg[0] == '\0')
        return fail("unknown kernel is 503 with a message");
    printf("argument: info=%d code=%d msg=\"%s\"\n", info, code, msg);

    if (ppmg_plan_destroy(plan) != 0) return fail("ppmg_plan_destroy");
    printf("ok\n");
    return 0;
}

/* field_up(lda, 1-g:Nm+g, 1-g:Nm+g, 1-g:Nm+g), Fortran order */
static size_t at(int c, int i, int j, int k, int lda, int g, int ext) {
    return (size_t)c + (size_t)lda * ((size_t)(i - 1 + g) + (size_t)ext * ((size_t)(j - 1 + g) + (size_t)ext * (size_t)(k - 1 + g)));
}

static int check_gpu(void) {
    const int N = 16, NM = N + 1, G = 2, LDA = 2, EXT = NM + 2 * G;
    const int np = N * N * N, nq = 1000;
    if (ppmg_init(3, PPMG_PREC_F64, -15, 0) != 0) return fail("ppmg_init");
    const double mn[3] = {0, 0, 0}, mx[3] = {1, 1, 1};
    const int bc[6] = {1, 1, 1, 1, 1, 1}, nm[3] = {NM, NM, NM}, g[3] = {G, G, G};
    int topo = 0, mesh = 0;
    if (ppmg_mktopo_field(3, mn, mx, bc, nm, g, &topo, &mesh) != 0) return fail("ppmg_mktopo_field");

    /* jittered lattice, one particle per cell, strengths in [-1, 1) */
    double* xp = malloc(sizeof(double) * 3 * (size_t)np);
    double* up = malloc(sizeof(double) * LDA * (size_t)np);
    double* xq = malloc(sizeof(double) * 3 * (size_t)nq);
    double* uq = malloc(sizeof(double) * LDA * (size_t)nq);
    const size_t nf = (size_t)LDA * EXT * EXT * EXT;
    double* field = malloc(sizeof(double) * nf);
    if (!xp || !up || !xq || !uq || !field) return fail("host allocation");
    double total[2] = {0, 0};
    for (int p = 0; p < np; ++p) {
        const int cell[3] = {p % N, (p / N) % N, p / (N * N)};
        for (int d = 0; d < 3; ++d) xp[3 * p + d] = (cell[d] + 0.2 + 0.6 * uniform01()) / N;
        for (int c = 0; c < LDA; ++c) {
            up[LDA * p + c] = 2.0 * uniform01() - 1.0;
            total[c] += up[LDA * p + c];
        }
    }
    for (int q = 0; q < nq; ++q)
        for (int d = 0; d < 3; ++d) xq[3 * q + d] = uniform01();

    const int kernels[2] = {PPMG_KERNEL_BSP2, PPMG_KERNEL_MP4};
    for (int kk = 0; kk < 2; ++kk) {
        const int kernel = kernels[kk];
        for (size_t i = 0; i < nf; ++i) field[i] = 0.0;
        if (ppmg_interp_p2m(xp, np, up, LDA, topo, mesh, kernel, g, field, PPMG_PREC_F64, PPMG_MEM_HOST) != 0)
            return fail("ppmg_interp_p2m");
        /* mass over the unique real nodes 1..N */
        for (int c = 0; c < LDA; ++c) {
            double s = 0.0;
            for (int k = 1; k <= N; ++k)
                for (int j = 1; j <= N; ++j)
                    for (int i = 1; i <= N; ++i) s += field[at(c, i, j, k, LDA, G, EXT)];
            const double scale = (double)np;   /* sum of |u| is at most np */
            if (!(fabs(s - total[c]) <= 1e-12 * scale)) return fail("p2m conserves mass");
            printf("kernel %d component %d: node sum %.15g, strength sum %.15g\n", kernel, c, s, total[c]);
        }
        /* the duplicate face node Nm equals node 1 on every axis */
        for (int k = 1; k <= NM; ++k)
            for (int j = 1; j <= NM; ++j)
                for (int c = 0; c < LDA; ++c) {
                    if (field[at(c, NM, j, k, LDA, G, EXT)] != field[at(c, 1, j, k, LDA, G, EXT)] ||
                        field[at(c, j, NM, k, LDA, G, EXT)] != field[at(c, j, 1, k, LDA, G, EXT)] ||
                        field[at(c, j, k, NM, LDA, G, EXT)] != field[at(c, j, k, 1, LDA, G, EXT)])
                        return fail("node Nm equals node 1 after the fold");
                }
        printf("kernel %d: duplicate face nodes equal node 1\n", kernel);
        /* M2P of a constant field */
        for (size_t i = 0; i < nf; ++i) field[i] = 2.5;
        if (ppmg_interp_m2p(xq, nq, uq, LDA, topo, mesh, kernel, g, field, PPMG_PREC_F64, PPMG_MEM_HOST) != 0)
            return fail("ppmg_interp_m2p");
        double worst = 0.0;
        for (int i = 0; i < nq * LDA; ++i) worst = fmax(worst, fabs(uq[i] - 2.5));
        if (!(worst <= 1e-13)) return fail("m2p of a constant field");
        printf("kernel %d: m2p of a constant field, max error %.3g\n", kernel, worst);
    }
    /* ghost_get of an integer field with a mask: every stored node holds the
     * code of the real node it images (coordinates wrapped into 1..N), ghost
     * nodes start at -7; afterwards a ghost node holds its real node's code
     * where the mask is set and still -7 where it is not */
    {
        const size_t nn = (size_t)EXT * EXT * EXT;
        int* fi = malloc(sizeof(int) * LDA * nn);
        int* mk = malloc(sizeof(int) * nn);
        if (!fi || !mk) return fail("host allocation");
#define WRAP(v) ((((v) - 1) % N + N) % N + 1)
#define CODE(i, j, k, c) (WRAP(i) + 100 * WRAP(j) + 10000 * WRAP(k) + 1000000 * (c))
        for (int k = 1 - G; k <= NM + G; ++k)
            for (int j = 1 - G; j <= NM + G; ++j)
                for (int i = 1 - G; i <= NM + G; ++i) {
                    const int real = i >= 1 && i <= NM && j >= 1 && j <= NM && k >= 1 && k <= NM;
                    for (int c = 0; c < LDA; ++c) fi[at(c, i, j, k, LDA, G, EXT)] = real ? CODE(i, j, k, c) : -7;
                    mk[at(0, i, j, k, 1, G, EXT)] = (i + 2 * j + 3 * k) % 3 != 0;
                }
        if (ppmg_map_field_ghost_kind_reg(fi, LDA, topo, mesh, g, PPMG_MAP_GHOST_GET, PPMG_KIND_I, mk, PPMG_MEM_HOST) != 0)
            return fail("ppmg_map_field_ghost_kind_reg");
        long set = 0, kept = 0;
        for (int k = 1 - G; k <= NM + G; ++k)
            for (int j = 1 - G; j <= NM + G; ++j)
                for (int i = 1 - G; i <= NM + G; ++i) {
                    const int real = i >= 1 && i <= NM && j >= 1 && j <= NM && k >= 1 && k <= NM;
                    const int on = mk[at(0, i, j, k, 1, G, EXT)];
                    for (int c = 0; c < LDA; ++c) {
                        const int want = real || on ? CODE(i, j, k, c) : -7;
                        if (fi[at(c, i, j, k, LDA, G, EXT)] != want) return fail("masked integer ghost_get");
                    }
                    if (!real) {
                        set += on;
                        kept += !on;
                    }
                }
#undef CODE
#undef WRAP
        printf("integer ghost_get with a mask: %ld ghost nodes set, %ld kept\n", set, kept);
        free(fi);
        free(mk);
    }
    if (ppmg_topo_destroy(topo) != 0) return fail("ppmg_topo_destroy");
    if (ppmg_finalize() != 0) return fail("ppmg_finalize");
    free(xp);
    free(up);
    free(xq);
    free(uq);
    free(field);
    printf("ok\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
        return 2;
    }
    if (strcmp(argv[1], "cpu") == 0) return check_cpu();
    if (strcmp(argv[1], "gpu") == 0) return check_gpu();
    fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
    return 2;
}
