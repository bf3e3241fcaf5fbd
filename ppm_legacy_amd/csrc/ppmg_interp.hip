// ppmg_interp.hip -- particle<->mesh kernels of libppmg (gfx950), general path.
//
// Every variant of the reference is covered here: 2D/3D, M'4/BSP2, any lda,
// FP64 / FP32-compensated / FP32, any stored node box.  P2M is a node-owned
// gather over the binned particles (no float atomics: every node sums its
// contributions in a fixed order, so results are bitwise reproducible);
// M2P and the adjoint are particle-owned gathers.  ppmg_p2m3d.hip holds the
// tiled fast path for the 3D M'4 workloads; this file is the reference-exact
// path it is checked against and the path every other variant runs on.
#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

static inline unsigned nblk(int64_t n, int b) {
    int64_t g = (n + b - 1) / b;
    return (unsigned)(g > 0 ? g : 1);
}

// ------------------------------------------------------------------------
// Per sorted particle: the KS 1D weights of every axis and the strengths,
// gathered into sorted order so a node's cell neighbourhood reads contiguous
// memory.  Weights either from the position (ppm_interp_p2m.f:800-853,
// 3077-3104, 3620-3651) or from the tables of ppm_rmsh_comp_weights
// (ppm_rmsh_remesh.f:618-866).
template <int DIM, int KS, typename PosT, typename ValT, bool TABLE>
__global__ void __launch_bounds__(256)
k_sorted_weights(const int* __restrict__ perm, const PosT* __restrict__ xp, const ValT* __restrict__ up,
                 int lda, int64_t np, DevMesh m, const ValT* __restrict__ wx1,
                 const ValT* __restrict__ wx2, const ValT* __restrict__ wx3, ValT* __restrict__ ws,
                 ValT* __restrict__ us) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    int p = perm[s];
    ValT* w = ws + s * (DIM * KS);
    if constexpr (TABLE) {
        const ValT* tab[3] = {wx1, wx2, wx3};
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int o = 0; o < KS; ++o) w[d * KS + o] = tab[d][(int64_t)p * KS + o];
    } else {
        const PosT* x = xp + (int64_t)p * DIM;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            int i;
            PosT frac = cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], i);
            ValT a[KS];
            weights_1d<KS, DIM, ValT>((ValT)frac, a);
#pragma unroll
            for (int o = 0; o < KS; ++o) w[d * KS + o] = a[o];
        }
    }
    for (int c = 0; c < lda; ++c) us[s * lda + c] = up[(int64_t)p * lda + c];
}

// A stencil cell of node n along axis d; false when it does not exist.
// one_period (a stored box holding the real nodes): the fold adds the images
// of a node one period away only (ppm_map_field_ghost_init.f:236-425, one
// block per face, edge and corner), so a cell more than one period from the
// node does not reach it -- which happens on a mesh of one cell per axis,
// where an M'4 stencil spans two periods.
__device__ __forceinline__ bool cell_of(const DevMesh& m, int d, bool wrap, bool one_period, int& c) {
    if (wrap) {
        const int N = m.N[d];
        const int c0 = c;
        c %= N;
        if (c < 0) c += N;
        return !one_period || abs(c - c0) <= N;
    }
    return c >= 0 && c < m.nbin[d];
}

// P2M gather: one thread per stored node.  Nodes real on every axis (0..N)
// sum their periodic images (the ghost fold of ppm_map_field_ghost_put: on a
// stored box the images one period away, in the compact layout all); nodes
// outside the real range keep the raw scatter, as in the reference.
template <int DIM, int KS, typename ValT, int LDA>
__global__ void __launch_bounds__(256)
k_p2m_gather(DevMesh m, FieldGeom f, ValT* __restrict__ field, const int* __restrict__ start,
             const ValT* __restrict__ ws, const ValT* __restrict__ us, int lda, int c0, int64_t nnodes) {
    using Acc = typename AccOf<ValT>::type;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnodes) return;
    int li = (int)(t % f.ext[0]);
    int64_t r = t / f.ext[0];
    int lj = (int)(r % f.ext[1]);
    int lk = (int)(r / f.ext[1]);
    int n[3] = {li + f.lo[0], lj + f.lo[1], DIM == 3 ? lk + f.lo[2] : 0};
    bool real = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        if (n[d] < 0 || n[d] > m.N[d]) real = false;
    constexpr int BASE = KS == 4 ? 1 : 0;
    Acc acc[LDA];
#pragma unroll
    for (int c = 0; c < LDA; ++c) acc[c].init();
    constexpr int KZ = DIM == 3 ? KS : 1;
    for (int oz = 0; oz < KZ; ++oz) {
        int cz = 0;
        if (DIM == 3) {
            cz = n[2] + BASE - oz;
            if (!cell_of(m, 2, real && m.periodic[2], f.swap[2], cz)) continue;
        }
        for (int oy = 0; oy < KS; ++oy) {
            int cy = n[1] + BASE - oy;
            if (!cell_of(m, 1, real && m.periodic[1], f.swap[1], cy)) continue;
            for (int ox = 0; ox < KS; ++ox) {
                int cx = n[0] + BASE - ox;
                if (!cell_of(m, 0, real && m.periodic[0], f.swap[0], cx)) continue;
                int key = cx + m.nbin[0] * (cy + m.nbin[1] * cz);
                int b = start[key], e = start[key + 1];
                for (int s = b; s < e; ++s) {
                    const ValT* w = ws + (int64_t)s * (DIM * KS);
                    ValT wt = w[ox] * w[KS + oy];
                    if (DIM == 3) wt = wt * w[2 * KS + oz];
                    const ValT* u = us + (int64_t)s * lda + c0;
#pragma unroll
                    for (int c = 0; c < LDA; ++c) acc[c].add(wt * u[c]);
                }
            }
        }
    }
    ValT* out = field + t * f.lda + c0;
#pragma unroll
    for (int c = 0; c < LDA; ++c) out[c] = (ValT)acc[c].value();
}

// Stored index of node v along axis d for a read, with ghost_get's periodic
// node map (ppm_map_field_ghost_get.f:257-280: node 0 and the ghosts below it
// take node v+N, node N and the ghosts above it take v-N); -1 if not stored.
// `filled` is false for a node that ghost_get leaves as stored (see
// real_on_open_axis): on a stored box it is then read at v itself.
__device__ __forceinline__ int read_index(const DevMesh& m, const FieldGeom& f, int d, int v, bool filled = true) {
    if (m.periodic[d]) {
        int N = m.N[d];
        if (f.swap[d]) {
            if (filled) {
                if (v <= 0) v += N;
                else if (v >= N) v -= N;
            }
        } else {
            v %= N;
            if (v < 0) v += N;
        }
    }
    int l = v - f.lo[d];
    return (l >= 0 && l < f.ext[d]) ? l : -1;
}

// M2P gather: one thread per particle (ppm_interp_m2p.f:895-1240, 3245-4106).
template <int DIM, int KS, typename PosT, typename ValT, int LDA>
__global__ void __launch_bounds__(256)
k_m2p_gather(DevMesh m, FieldGeom f, const ValT* __restrict__ field, const PosT* __restrict__ xp,
             ValT* __restrict__ up, int lda, int c0, int64_t np, int* __restrict__ err) {
    using Acc = typename AccOf<ValT>::type;
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    constexpr int BASE = KS == 4 ? 1 : 0;
    const PosT* x = xp + p * DIM;
    // A position that is not a number, or so far away that its node index
    // does not fit an int, reads no node: it is counted with the particles
    // whose stencil leaves the stored field, and its values are NaN.
    bool sane = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const PosT x0 = scaled_coord<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d]);
        sane = sane && x0 > PosT(-1.0e9) && x0 < PosT(1.0e9);
    }
    if (!sane) {
        atomicAdd(err, 1);
        ValT* u = up + p * lda + c0;
#pragma unroll
        for (int c = 0; c < LDA; ++c) u[c] = (ValT)__builtin_nan("");
        return;
    }
    ValT a[3][KS];
    int nb[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int i;
        PosT frac = cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], i);
        weights_1d<KS, DIM, ValT>((ValT)frac, a[d]);
        nb[d] = i - BASE;
    }
    constexpr int KZ = DIM == 3 ? KS : 1;
    Acc acc[KZ][LDA];   // one partial sum per z offset: independent FMA chains
#pragma unroll
    for (int oz = 0; oz < KZ; ++oz)
#pragma unroll
        for (int c = 0; c < LDA; ++c) acc[oz][c].init();
    bool bad = false;
#pragma unroll
    for (int ox = 0; ox < KS; ++ox) {
        const int vx = nb[0] + ox;
        const bool rx = real_on_open_axis(m, 0, vx);
#pragma unroll
        for (int oy = 0; oy < KS; ++oy) {
            const int vy = nb[1] + oy;
            const bool rxy = rx && real_on_open_axis(m, 1, vy);
#pragma unroll
            for (int oz = 0; oz < KZ; ++oz) {
                const int vz = nb[2] + oz;
                const bool filled = rxy && (DIM == 2 || real_on_open_axis(m, 2, vz));
                const int ix = read_index(m, f, 0, vx, filled);
                const int iy = read_index(m, f, 1, vy, filled);
                const int iz = DIM == 3 ? read_index(m, f, 2, vz, filled) : 0;
                if (ix < 0 || iy < 0 || iz < 0) {
                    bad = true;
                    continue;
                }
                ValT w = a[0][ox] * a[1][oy];
                if (DIM == 3) w = w * a[2][oz];
                const ValT* fv = field + ((int64_t)ix + (int64_t)f.ext[0] * (iy + (int64_t)f.ext[1] * iz)) * f.lda + c0;
#pragma unroll
                for (int c = 0; c < LDA; ++c) acc[oz][c].add(w * fv[c]);
            }
        }
    }
    if (bad) atomicAdd(err, 1);
#pragma unroll
    for (int oz = 1; oz < KZ; ++oz)
#pragma unroll
        for (int c = 0; c < LDA; ++c) acc[0][c].merge(acc[oz][c]);
    ValT* u = up + p * lda + c0;
#pragma unroll
    for (int c = 0; c < LDA; ++c) u[c] = (ValT)acc[0][c].value();
}

// Stored indices along axis d of the real nodes that receive stencil node v
// in the forward P2M: on a periodic axis its image r = v mod N and, for r = 0,
// node N (both real, both given every periodic image's sum; on a stored box
// only those one period from v); on an open axis v itself if it is a real
// node 0..N.  Indices outside the box are dropped.
__device__ __forceinline__ int real_images(const DevMesh& m, const FieldGeom& f, int d, int v, int (&out)[2]) {
    int n = 0;
    if (m.periodic[d]) {
        const int N = m.N[d];
        int r = v % N;
        if (r < 0) r += N;
        // a stored box folds images one period away only (cell_of)
        const bool one = f.swap[d];
        int l = r - f.lo[d];
        if (l >= 0 && l < f.ext[d] && (!one || abs(v - r) <= N)) out[n++] = l;
        l = N - f.lo[d];
        if (r == 0 && l >= 0 && l < f.ext[d] && (!one || abs(v - N) <= N)) out[n++] = l;
    } else if (v >= 0 && v <= m.N[d]) {
        const int l = v - f.lo[d];
        if (l >= 0 && l < f.ext[d]) out[n++] = l;
    }
    return n;
}

// Adjoint of P2M (no reference counterpart), one thread per particle in the
// forward's sorted order, reusing its permutation and cell keys:
//   dL/du_p = sum_n W(x0_p - n) g_n,  dL/dx_p,d = sum_n (g_n . u_p) dW/dx_d,
// where n runs over the stored nodes the forward writes particle p into: for
// each stencil node v, every real node congruent to v on the periodic axes
// (v on the open ones), and v itself where it is a stored node outside the
// real range (a ghost node keeps the raw sum).
template <int DIM, int KS, typename PosT, typename ValT, int LDA>
__global__ void __launch_bounds__(256)
k_adjoint(DevMesh m, FieldGeom f, const ValT* __restrict__ g, const PosT* __restrict__ xp,
          const ValT* __restrict__ up, const int* __restrict__ perm, const int* __restrict__ key, int lda,
          int c0, int64_t np, ValT* __restrict__ grad_u, ValT* __restrict__ grad_x, int accumulate_x) {
    using Acc = typename AccOf<ValT>::type;
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    constexpr int BASE = KS == 4 ? 1 : 0;
    int p = perm[s];
    int k = key[p];
    int cell[3] = {k % m.nbin[0], (k / m.nbin[0]) % m.nbin[1], DIM == 3 ? k / (m.nbin[0] * m.nbin[1]) : 0};
    const PosT* x = xp + (int64_t)p * DIM;
    ValT a[3][KS], da[3][KS];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int i;
        ValT frac = (ValT)cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], i);
        weights_1d<KS, DIM, ValT>(frac, a[d]);
        if constexpr (KS == 4) {
            mp4_unrolled_deriv<ValT>(frac, da[d]);
        } else {
            da[d][0] = ValT(-1);
            da[d][1] = ValT(1);
        }
        ValT sc = (ValT)m.dxi[d];
#pragma unroll
        for (int o = 0; o < KS; ++o) da[d][o] = da[d][o] * sc;
    }
    if (DIM == 2) {
#pragma unroll
        for (int o = 0; o < KS; ++o) { a[2][o] = o == 0 ? ValT(1) : ValT(0); da[2][o] = ValT(0); }
    }
    ValT uu[LDA];
#pragma unroll
    for (int c = 0; c < LDA; ++c) uu[c] = up[(int64_t)p * lda + c0 + c];
    Acc gu[LDA], gx[3];
#pragma unroll
    for (int c = 0; c < LDA; ++c) gu[c].init();
#pragma unroll
    for (int d = 0; d < 3; ++d) gx[d].init();
    constexpr int KZ = DIM == 3 ? KS : 1;
    for (int oz = 0; oz < KZ; ++oz) {
        const int vz = cell[2] - BASE + oz;
        int iz[2] = {0, 0};
        const int nz = DIM == 3 ? real_images(m, f, 2, vz, iz) : 1;
        const bool rz = DIM == 2 || (vz >= 0 && vz <= m.N[2]);
        const int lz = DIM == 3 ? vz - f.lo[2] : 0;
        for (int oy = 0; oy < KS; ++oy) {
            const int vy = cell[1] - BASE + oy;
            int iy[2] = {0, 0};
            const int ny = real_images(m, f, 1, vy, iy);
            const bool ry = rz && vy >= 0 && vy <= m.N[1];
            const int ly = vy - f.lo[1];
            for (int ox = 0; ox < KS; ++ox) {
                const int vx = cell[0] - BASE + ox;
                int ix[2] = {0, 0};
                const int nx = real_images(m, f, 0, vx, ix);
                const int lx = vx - f.lo[0];
                ValT gs[LDA];
#pragma unroll
                for (int c = 0; c < LDA; ++c) gs[c] = ValT(0);
                bool any = false;
                for (int a2 = 0; a2 < nz; ++a2)
                    for (int a1 = 0; a1 < ny; ++a1)
                        for (int a0 = 0; a0 < nx; ++a0) {
                            const ValT* gn =
                                g + ((int64_t)ix[a0] + (int64_t)f.ext[0] * (iy[a1] + (int64_t)f.ext[1] * iz[a2])) * f.lda + c0;
#pragma unroll
                            for (int c = 0; c < LDA; ++c) gs[c] += gn[c];
                            any = true;
                        }
                // v outside the real range: the ghost node v, if stored, holds its raw sum
                if (!(ry && vx >= 0 && vx <= m.N[0]) && lx >= 0 && lx < f.ext[0] && ly >= 0 && ly < f.ext[1] &&
                    lz >= 0 && lz < f.ext[2]) {
                    const ValT* gn = g + ((int64_t)lx + (int64_t)f.ext[0] * (ly + (int64_t)f.ext[1] * lz)) * f.lda + c0;
#pragma unroll
                    for (int c = 0; c < LDA; ++c) gs[c] += gn[c];
                    any = true;
                }
                if (!any) continue;   // p writes no stored node here: zero cotangent
                ValT w = a[0][ox] * a[1][oy] * a[2][oz];
                ValT gdu = ValT(0);
#pragma unroll
                for (int c = 0; c < LDA; ++c) {
                    gu[c].add(w * gs[c]);
                    gdu += gs[c] * uu[c];
                }
                gx[0].add(gdu * (da[0][ox] * a[1][oy] * a[2][oz]));
                gx[1].add(gdu * (a[0][ox] * da[1][oy] * a[2][oz]));
                if (DIM == 3) gx[2].add(gdu * (a[0][ox] * a[1][oy] * da[2][oz]));
            }
        }
    }
    if (grad_u) {
#pragma unroll
        for (int c = 0; c < LDA; ++c) grad_u[(int64_t)p * lda + c0 + c] = (ValT)gu[c].value();
    }
    if (grad_x) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            ValT v = (ValT)gx[d].value();
            if (accumulate_x) v += grad_x[(int64_t)p * DIM + d];
            grad_x[(int64_t)p * DIM + d] = v;
        }
    }
}

// ppm_rmsh_comp_weights (ppm_rmsh_comp_weights.f:853-994): weights by
// |node - x0| with <= branches, zero outside the support.
template <int DIM, int KS, typename PosT, typename ValT>
__global__ void __launch_bounds__(256)
k_comp_weights(const PosT* __restrict__ xp, int64_t np, DevMesh m, ValT* __restrict__ wx1,
               ValT* __restrict__ wx2, ValT* __restrict__ wx3) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    ValT* tab[3] = {wx1, wx2, wx3};
    const PosT* x = xp + p * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        PosT x0 = scaled_coord<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d]);
        int ip = (int)x0 + 1;                       // INT(..) + 2 - istart, istart = 1
        int lo = KS == 4 ? ip - 1 : ip;
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            PosT r = (PosT)(lo + q - 1) - x0;
            r = r < PosT(0) ? -r : r;
            PosT w = 0;
            if (KS == 2) {
                if (r <= PosT(1)) w = PosT(1) - r;
            } else if (r <= PosT(2)) {
                if (r <= PosT(1)) w = PosT(1) - (r * r) * (PosT(2.5) - PosT(1.5) * r);
                else w = PosT(2) + (PosT(-4) + (PosT(2.5) - PosT(0.5) * r) * r) * r;
            }
            tab[d][p * KS + q] = (ValT)w;
        }
    }
}

// ppm_impose_part_bc (ppm_impose_part_bc.f:173-199): two separate IFs.
template <int DIM, typename PosT>
__global__ void __launch_bounds__(256) k_impose_bc(PosT* __restrict__ xp, int64_t np, DevMesh m) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (!m.periodic[d]) continue;
        PosT mn = (PosT)m.minv[d], mx = (PosT)m.maxv[d];
        PosT len = mx - mn;
        PosT v = xp[p * DIM + d];
        if (v < mn) v = v + len;
        if (v >= mx) v = v - len;
        xp[p * DIM + d] = v;
    }
}

// ppm_rmsh_create_part (ppm_rmsh_create_part.f:427-681): band test per node.
template <typename ValT>
__device__ __forceinline__ bool cp_in_band(const ValT* fv, int lda, const double* cw, ValT lo, ValT hi) {
    ValT s;
    if (lda == 1 && !cw) {
        s = fv[0] < ValT(0) ? -fv[0] : fv[0];
    } else {
        ValT acc = 0;
        for (int c = 0; c < lda; ++c) {
            ValT v = cw ? fv[c] * (ValT)cw[c] : fv[c];
            acc = acc + v * v;
        }
        s = sqrt(acc);
    }
    return s > lo && s < hi;
}

struct CPArgs {
    int nn[3];
    double lo, hi;
    double cw[16];
    int use_cw;
    int resetpos;
    int64_t maxnp;
};

template <typename ValT>
__global__ void __launch_bounds__(256)
k_cp_flag(FieldGeom f, const ValT* __restrict__ field, CPArgs a, int64_t nn, int* __restrict__ flag) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % a.nn[0]);
    int64_t r = t / a.nn[0];
    int j = (int)(r % a.nn[1]);
    int k = (int)(r / a.nn[1]);
    int64_t off = ((int64_t)(i - f.lo[0]) + (int64_t)f.ext[0] * ((j - f.lo[1]) + (int64_t)f.ext[1] * (k - f.lo[2]))) * f.lda;
    flag[t] = cp_in_band<ValT>(field + off, f.lda, a.use_cw ? a.cw : nullptr, (ValT)a.lo, (ValT)a.hi) ? 1 : 0;
}

template <typename PosT, typename ValT>
__global__ void __launch_bounds__(256)
k_cp_write(DevMesh m, FieldGeom f, const ValT* __restrict__ field, CPArgs a, int64_t nn,
           const int* __restrict__ flag, const int* __restrict__ pos, PosT* __restrict__ xp,
           ValT* __restrict__ up, FieldGeom fw, const ValT* __restrict__ wfield, ValT* __restrict__ wp) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn || !flag[t]) return;
    int64_t q = pos[t];
    if (q >= a.maxnp) return;
    int idx[3];
    idx[0] = (int)(t % a.nn[0]);
    int64_t r = t / a.nn[0];
    idx[1] = (int)(r % a.nn[1]);
    idx[2] = (int)(r / a.nn[1]);
    int64_t off = ((int64_t)(idx[0] - f.lo[0]) + (int64_t)f.ext[0] * ((idx[1] - f.lo[1]) + (int64_t)f.ext[1] * (idx[2] - f.lo[2]))) * f.lda;
    for (int c = 0; c < f.lda; ++c) up[q * f.lda + c] = field[off + c];
    if (wfield) {
        // slave field: its values at the same node (ppm_rmsh_create_part.f:626-632)
        int64_t woff = ((int64_t)(idx[0] - fw.lo[0]) +
                        (int64_t)fw.ext[0] * ((idx[1] - fw.lo[1]) + (int64_t)fw.ext[1] * (idx[2] - fw.lo[2]))) * fw.lda;
        for (int c = 0; c < fw.lda; ++c) wp[q * fw.lda + c] = wfield[woff + c];
    }
    if (a.resetpos) {
        for (int d = 0; d < m.dim; ++d) {
            PosT dx = ((PosT)m.maxv[d] - (PosT)m.minv[d]) / (PosT)m.N[d];
            xp[q * m.dim + d] = (PosT)idx[d] * dx + (PosT)m.minv[d];
        }
    }
}

// ================================================================ host side
template <typename PosT, typename ValT>
static int p2m_typed(ppmg_plan* plan, const DevMesh& m, int kernel, int64_t np, const PosT* xp, int lda,
                     const ValT* up, const FieldGeom& fg, ValT* field, const ValT* const* wtab, hipStream_t s) {
    const int KS = kernel == PPMG_KERNEL_MP4 ? 4 : 2;
    const int dim = m.dim;
    size_t wbytes = sizeof(ValT) * (size_t)(np > 0 ? np : 1) * dim * KS;
    size_t ubytes = sizeof(ValT) * (size_t)(np > 0 ? np : 1) * lda;
    if (plan->wsorted.ensure(wbytes) || plan->usorted.ensure(ubytes)) return PPMG_ERROR_FATAL;
    ValT* ws = plan->wsorted.as<ValT>();
    ValT* us = plan->usorted.as<ValT>();
    if (materialize_perm(plan, s)) return PPMG_ERROR_FATAL;
    const int* perm = plan->perm.as<int>();
    const ValT* t1 = wtab ? wtab[0] : nullptr;
    const ValT* t2 = wtab ? wtab[1] : nullptr;
    const ValT* t3 = wtab ? wtab[2] : nullptr;
    if (np > 0) {
        unsigned g = nblk(np, 256);
#define SW(D, K, T) hipLaunchKernelGGL((k_sorted_weights<D, K, PosT, ValT, T>), dim3(g), dim3(256), 0, s, perm, xp, up, lda, np, m, t1, t2, t3, ws, us)
        if (dim == 3 && KS == 4) { if (wtab) SW(3, 4, true); else SW(3, 4, false); }
        else if (dim == 3) { if (wtab) SW(3, 2, true); else SW(3, 2, false); }
        else if (KS == 4) { if (wtab) SW(2, 4, true); else SW(2, 4, false); }
        else { if (wtab) SW(2, 2, true); else SW(2, 2, false); }
#undef SW
        PPMG_HIP(hipGetLastError());
    }
    mark(plan, s);
    int64_t nnodes = (int64_t)fg.ext[0] * fg.ext[1] * fg.ext[2];
    const int* start = plan->start.as<int>();
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
        unsigned g = nblk(nnodes, 256);
#define GA(D, K, LL) hipLaunchKernelGGL((k_p2m_gather<D, K, ValT, LL>), dim3(g), dim3(256), 0, s, m, fg, field, start, ws, us, lda, c0, nnodes)
#define GL(D, K) do { if (L == 1) GA(D, K, 1); else if (L == 2) GA(D, K, 2); else if (L == 3) GA(D, K, 3); else GA(D, K, 4); } while (0)
        if (dim == 3 && KS == 4) GL(3, 4);
        else if (dim == 3) GL(3, 2);
        else if (KS == 4) GL(2, 4);
        else GL(2, 2);
#undef GL
#undef GA
        PPMG_HIP(hipGetLastError());
    }
    mark(plan, s);
    return 0;
}

int p2m_general(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, int lda,
                const void* up, const FieldGeom& fg, void* field, const void* const* wtab, hipStream_t s) {
    if (prec == PPMG_PREC_F64)
        return p2m_typed<double, double>(plan, m, kernel, np, (const double*)xp, lda, (const double*)up, fg,
                                         (double*)field, (const double* const*)wtab, s);
    if (prec == PPMG_PREC_F32C)
        return p2m_typed<double, float>(plan, m, kernel, np, (const double*)xp, lda, (const float*)up, fg,
                                        (float*)field, (const float* const*)wtab, s);
    return p2m_typed<float, float>(plan, m, kernel, np, (const float*)xp, lda, (const float*)up, fg,
                                   (float*)field, (const float* const*)wtab, s);
}

template <typename PosT, typename ValT>
static int m2p_typed(const DevMesh& m, int kernel, int64_t np, const PosT* xp, int lda, const FieldGeom& fg,
                     const ValT* field, ValT* up, int* err, hipStream_t s) {
    const int KS = kernel == PPMG_KERNEL_MP4 ? 4 : 2;
    if (np <= 0) return 0;
    unsigned g = nblk(np, 256);
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
#define GA(D, K, LL) hipLaunchKernelGGL((k_m2p_gather<D, K, PosT, ValT, LL>), dim3(g), dim3(256), 0, s, m, fg, field, xp, up, lda, c0, np, err)
#define GL(D, K) do { if (L == 1) GA(D, K, 1); else if (L == 2) GA(D, K, 2); else if (L == 3) GA(D, K, 3); else GA(D, K, 4); } while (0)
        if (m.dim == 3 && KS == 4) GL(3, 4);
        else if (m.dim == 3) GL(3, 2);
        else if (KS == 4) GL(2, 4);
        else GL(2, 2);
#undef GL
#undef GA
        PPMG_HIP(hipGetLastError());
    }
    return 0;
}

int m2p_general(const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, int lda, const FieldGeom& fg,
                const void* field, void* up, int* err, hipStream_t s) {
    if (prec == PPMG_PREC_F64)
        return m2p_typed<double, double>(m, kernel, np, (const double*)xp, lda, fg, (const double*)field,
                                         (double*)up, err, s);
    if (prec == PPMG_PREC_F32C)
        return m2p_typed<double, float>(m, kernel, np, (const double*)xp, lda, fg, (const float*)field,
                                        (float*)up, err, s);
    return m2p_typed<float, float>(m, kernel, np, (const float*)xp, lda, fg, (const float*)field, (float*)up, err, s);
}

template <typename PosT, typename ValT>
static int adjoint_typed(ppmg_plan* plan, const DevMesh& m, int kernel, int64_t np, const PosT* xp, int lda,
                         const ValT* up, const FieldGeom& fg, const ValT* g, ValT* gu, ValT* gx, hipStream_t s) {
    const int KS = kernel == PPMG_KERNEL_MP4 ? 4 : 2;
    if (np <= 0) return 0;
    unsigned gr = nblk(np, 256);
    if (materialize_perm(plan, s) || materialize_keys(plan, s)) return PPMG_ERROR_FATAL;
    const int* perm = plan->perm.as<int>();
    const int* key = plan->key.as<int>();
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
        int acc = c0 > 0 ? 1 : 0;
#define GA(D, K, LL) hipLaunchKernelGGL((k_adjoint<D, K, PosT, ValT, LL>), dim3(gr), dim3(256), 0, s, m, fg, g, xp, up, perm, key, lda, c0, np, gu, gx, acc)
#define GL(D, K) do { if (L == 1) GA(D, K, 1); else if (L == 2) GA(D, K, 2); else if (L == 3) GA(D, K, 3); else GA(D, K, 4); } while (0)
        if (m.dim == 3 && KS == 4) GL(3, 4);
        else if (m.dim == 3) GL(3, 2);
        else if (KS == 4) GL(2, 4);
        else GL(2, 2);
#undef GL
#undef GA
        PPMG_HIP(hipGetLastError());
    }
    return 0;
}

int adjoint_general(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, int lda,
                    const void* up, const FieldGeom& fg, const void* g, void* gu, void* gx, hipStream_t s) {
    if (prec == PPMG_PREC_F64)
        return adjoint_typed<double, double>(plan, m, kernel, np, (const double*)xp, lda, (const double*)up, fg,
                                             (const double*)g, (double*)gu, (double*)gx, s);
    if (prec == PPMG_PREC_F32C)
        return adjoint_typed<double, float>(plan, m, kernel, np, (const double*)xp, lda, (const float*)up, fg,
                                            (const float*)g, (float*)gu, (float*)gx, s);
    return adjoint_typed<float, float>(plan, m, kernel, np, (const float*)xp, lda, (const float*)up, fg,
                                       (const float*)g, (float*)gu, (float*)gx, s);
}

int comp_weights_general(const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, void* w1, void* w2,
                         void* w3, hipStream_t s) {
    if (np <= 0) return 0;
    unsigned g = nblk(np, 256);
    const bool mp4 = kernel == PPMG_KERNEL_MP4;
#define CW(D, K, P, V) hipLaunchKernelGGL((k_comp_weights<D, K, P, V>), dim3(g), dim3(256), 0, s, (const P*)xp, np, m, (V*)w1, (V*)w2, (V*)w3)
#define CWP(D, K) do { if (prec == PPMG_PREC_F64) CW(D, K, double, double); else if (prec == PPMG_PREC_F32C) CW(D, K, double, float); else CW(D, K, float, float); } while (0)
    if (m.dim == 3) { if (mp4) CWP(3, 4); else CWP(3, 2); }
    else { if (mp4) CWP(2, 4); else CWP(2, 2); }
#undef CWP
#undef CW
    PPMG_HIP(hipGetLastError());
    return 0;
}

int impose_bc_general(const DevMesh& m, int prec, int64_t np, void* xp, hipStream_t s) {
    if (np <= 0) return 0;
    unsigned g = nblk(np, 256);
    if (m.dim == 3) {
        if (prec == PPMG_PREC_F32) hipLaunchKernelGGL((k_impose_bc<3, float>), dim3(g), dim3(256), 0, s, (float*)xp, np, m);
        else hipLaunchKernelGGL((k_impose_bc<3, double>), dim3(g), dim3(256), 0, s, (double*)xp, np, m);
    } else {
        if (prec == PPMG_PREC_F32) hipLaunchKernelGGL((k_impose_bc<2, float>), dim3(g), dim3(256), 0, s, (float*)xp, np, m);
        else hipLaunchKernelGGL((k_impose_bc<2, double>), dim3(g), dim3(256), 0, s, (double*)xp, np, m);
    }
    PPMG_HIP(hipGetLastError());
    return 0;
}

template <typename PosT, typename ValT>
static int create_part_typed(ppmg_plan* plan, const DevMesh& m, const FieldGeom& fg, const ValT* field,
                             const CPArgs& a, int64_t* np_out, PosT* xp, ValT* up, const FieldGeom& fw,
                             const ValT* wfield, ValT* wp, hipStream_t s) {
    int64_t nn = (int64_t)a.nn[0] * a.nn[1] * a.nn[2];
    if (plan->misc.ensure(sizeof(int) * (size_t)(nn + 1)) || plan->misc2.ensure(sizeof(int) * (size_t)(nn + 1)))
        return PPMG_ERROR_FATAL;
    int* flag = plan->misc.as<int>();
    int* pos = plan->misc2.as<int>();
    unsigned g = nblk(nn, 256);
    hipLaunchKernelGGL((k_cp_flag<ValT>), dim3(g), dim3(256), 0, s, fg, field, a, nn, flag);
    PPMG_HIP(hipGetLastError());
    if (exclusive_scan(flag, nn, pos, plan->scan_tmp, 0, s)) return PPMG_ERROR_FATAL;
    // total = pos[nn-1] + flag[nn-1]
    int last[2] = {0, 0};
    if (nn > 0) {
        PPMG_HIP(hipMemcpyAsync(&last[0], pos + nn - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        PPMG_HIP(hipMemcpyAsync(&last[1], flag + nn - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    hipLaunchKernelGGL((k_cp_write<PosT, ValT>), dim3(g), dim3(256), 0, s, m, fg, field, a, nn, flag, pos, xp, up, fw,
                       wfield, wp);
    PPMG_HIP(hipGetLastError());
    PPMG_HIP(hipStreamSynchronize(s));
    *np_out = (int64_t)last[0] + last[1];
    return 0;
}

int create_part_general(ppmg_plan* plan, const DevMesh& m, int prec, const FieldGeom& fg, const void* field,
                        const double* cutoff, const double* cw, int resetpos, int64_t maxnp, int64_t* np_out,
                        void* xp, void* up, const FieldGeom* fw, const void* wfield, void* wp, hipStream_t s) {
    FieldGeom fws = fw ? *fw : fg;
    if (!fw) wfield = nullptr;
    CPArgs a{};
    for (int d = 0; d < 3; ++d) a.nn[d] = 1;
    for (int d = 0; d < m.dim; ++d) a.nn[d] = m.upper_open[d] ? m.N[d] : m.N[d] + 1;
    a.lo = cutoff[0];
    a.hi = cutoff[1];
    a.use_cw = cw != nullptr;
    if (cw)
        for (int c = 0; c < fg.lda && c < 16; ++c) a.cw[c] = cw[c];
    a.resetpos = resetpos;
    a.maxnp = maxnp;
    if (prec == PPMG_PREC_F64)
        return create_part_typed<double, double>(plan, m, fg, (const double*)field, a, np_out, (double*)xp, (double*)up,
                                                 fws, (const double*)wfield, (double*)wp, s);
    if (prec == PPMG_PREC_F32C)
        return create_part_typed<double, float>(plan, m, fg, (const float*)field, a, np_out, (double*)xp, (float*)up,
                                                fws, (const float*)wfield, (float*)wp, s);
    return create_part_typed<float, float>(plan, m, fg, (const float*)field, a, np_out, (float*)xp, (float*)up, fws,
                                           (const float*)wfield, (float*)wp, s);
}

}  // namespace ppmg
