// ppmg_adj3d.hip -- adjoint of the 3D M'4 P2M (no reference counterpart),
// tiled for gfx950 on the structure of the M2P gather (ppmg_m2p3d.hip).
//
// For a loss L of the P2M field with cotangent g on the unique nodes:
//   dL/du_p,c = sum_n W(x_n - x_p) g_n,c                     (M2P of g)
//   dL/dx_p,d = sum_c u_p,c sum_n dW/dx_p,d (x_n - x_p) g_n,c
// with dW/dx_d the M'4 derivative stencil along d times dxi_d.  The forward's
// binning (permutation, cell heads) is reused.
//
// Layout and lane mapping as the M2P tile: a workgroup (4 wavefronts, 16
// rows of 16 lanes) owns 13 x 16 cells in x,y and marches a z chunk; an LDS
// ring holds 4 planes of g (16 x 19 nodes, component-major).  Lane ex of row
// `row` reads the 4x4 (y,z) column of node column tx0-1+ex and contracts it,
// for the particles of lanes ex+1, ex, ex-1, ex-2 (k = 0..3), three ways:
//   C0_k = sum wy wz g,   Cy_k = sum dwy wz g,   Cz_k = sum wy dwz g
// (the y and z weights of those particles arrive by DPP row shifts); the
// partial sums return to their particles, where the x weights finish them:
//   dL/du_c = sum_k wx_k C0_k,         dL/dx_0 = sum_c u_c sum_k dwx_k C0_k,
//   dL/dx_1 = sum_c u_c sum_k wx_k Cy_k, dL/dx_2 = sum_c u_c sum_k wx_k Cz_k.
// All sums are FP64 in registers; the FP32 forms keep FP32 only for g, u and
// the outputs.
#include <atomic>

#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

constexpr int A3_TX = 13;
constexpr int A3_TY = 16;
constexpr int A3_RY = A3_TY + 3;
constexpr int A3_PL = A3_RY * 16;
constexpr int A3_THREADS = 16 * A3_TY;

// Cotangent of stencil node u (unwrapped, periodic axes) when the forward
// wrote a stored box: the forward's value at a stored node real on every axis
// is the folded sum of the unique node wrap(u), copied to node N of a boxed
// axis, and a ghost node holds the raw, unwrapped sum.  So the particle
// collects g over the stored copies of wrap(u) (node 0 and its copy N on each
// boxed axis) and, when u itself is a stored ghost node, g at u.  Summed in
// FP64 (component c of the stored box).
template <typename ValT>
__device__ __forceinline__ double box_cotangent(const DevMesh& m, const FieldGeom& fg, const ValT* __restrict__ g,
                                                int c, const int (&u)[3]) {
    int s0[3], s1[3];
    bool real = true, stored = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int N = m.N[d], hi = fg.lo[d] + fg.ext[d] - 1;
        // u lies a few periods at most from 0..N-1 (one for N >= 16): wrap
        // by steps rather than an integer division
        int w = u[d];
        while (w < 0) w += N;
        while (w >= N) w -= N;
        s0[d] = w;
        s1[d] = w == 0 && hi >= N ? N : -1;
        if (u[d] < 0 || u[d] > N) real = false;
        if (u[d] < fg.lo[d] || u[d] > hi) stored = false;
    }
    auto at = [&](int x, int y, int z) {
        return (double)g[(((int64_t)(z - fg.lo[2]) * fg.ext[1] + (y - fg.lo[1])) * fg.ext[0] + (x - fg.lo[0])) *
                             fg.lda + c];
    };
    double acc = 0.0;
    for (int a = 0; a < 2; ++a) {
        const int x = a ? s1[0] : s0[0];
        if (x < 0) continue;
        for (int b = 0; b < 2; ++b) {
            const int y = b ? s1[1] : s0[1];
            if (y < 0) continue;
            for (int e = 0; e < 2; ++e) {
                const int z = e ? s1[2] : s0[2];
                if (z >= 0) acc += at(x, y, z);
            }
        }
    }
    if (!real && stored) acc += at(u[0], u[1], u[2]);
    return acc;
}

// BOXED: the cotangent is on a stored box (some axis holds 0..N with its
// ghosts, the reference's field_up(1-g:Nm+g)); the ring fetch reads it
// through box_cotangent.  Otherwise it is the compact layout.
// TABLE: the forward's binning is a slot table (byte counts, particle r of
// cell c at bt.table[r * ncells + c]) instead of heads and a permutation.
template <int LDA, typename PosT, typename ValT, typename AccT, bool BOXED, bool TABLE>
__global__ void __launch_bounds__(A3_THREADS) __attribute__((amdgpu_waves_per_eu(3)))
k_a3_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, BinTable bt,
          const PosT* __restrict__ xp, const ValT* __restrict__ up, const ValT* __restrict__ g, FieldGeom fg,
          ValT* __restrict__ grad_u, ValT* __restrict__ grad_x, int lda, int c0, int tz, int accumulate_x) {
    __shared__ ValT ring[4 * LDA * A3_PL];   // [plane][c][ry][nx]
    const int tid = threadIdx.x;
    const int ex = tid & 15, row = tid >> 4;
    const int Nx = m.N[0], Ny = m.N[1], Nz = m.N[2];
    const int ntx = (Nx + A3_TX - 1) / A3_TX, nty = (Ny + A3_TY - 1) / A3_TY;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int tx0 = (int)(tile % ntx) * A3_TX, ty0 = (int)((tile / ntx) % nty) * A3_TY;
    const int z0 = (int)(tile / ((int64_t)ntx * nty)) * tz;
    const int z1 = min(z0 + tz, Nz);
    const int nxo = min(A3_TX, Nx - tx0), nyo = min(A3_TY, Ny - ty0);
    const int64_t plane = (int64_t)Nx * Ny;

    // A thread stages whole nodes (all LDA components, adjacent in the field):
    // the ring stores of one component then write consecutive nodes.
    constexpr int NN = (A3_PL + A3_THREADS - 1) / A3_THREADS;   // nodes per thread and plane
    // BOXED and box_pipe: the tile's ring window lies inside the real range
    // in x and y (0 < u < N) and a box plane fits 32-bit byte offsets; its
    // planes inside the real range in z are then plain stored nodes (no other
    // copy) and take the same staged fetch as the compact layout.  Other
    // planes of a BOXED tile are read through box_cotangent between the
    // barriers.
    const bool box_pipe = BOXED && tx0 - 1 > 0 && tx0 + 14 < Nx && ty0 - 1 > 0 && ty0 + A3_RY - 2 < Ny &&
                          (int64_t)fg.ext[0] * fg.ext[1] * fg.lda * (int64_t)sizeof(ValT) < ((int64_t)1 << 31);
    auto staged = [&](int n) { return !BOXED || (box_pipe && n > 0 && n < Nz); };
    ValT stage[NN][LDA];
    // Byte offset within a field plane of each node this thread stages (the
    // same in every plane; a compact plane fits in 2^31 bytes, checked at
    // launch).
    unsigned soff[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        const int node = tid + j * A3_THREADS;
        const int nx = node & 15, ry = node >> 4;
        if constexpr (BOXED) {
            soff[j] = box_pipe ? (unsigned)((((int64_t)(ty0 - 1 + ry - fg.lo[1]) * fg.ext[0] + (tx0 - 1 + nx - fg.lo[0])) *
                                                 fg.lda + c0) * (int64_t)sizeof(ValT))
                               : 0u;
        } else {
            int gx = (tx0 - 1 + nx) % Nx;
            if (gx < 0) gx += Nx;
            int gy = (ty0 - 1 + ry) % Ny;
            if (gy < 0) gy += Ny;
            soff[j] = (unsigned)(((gy * Nx + gx) * lda + c0) * (int)sizeof(ValT));
        }
    }
    auto fetch_plane = [&](int n) {
        const char* base;
        if constexpr (BOXED) {
            base = (const char*)(g + (int64_t)(n - fg.lo[2]) * fg.ext[0] * fg.ext[1] * fg.lda);
        } else {
            int nz = n % Nz;
            if (nz < 0) nz += Nz;
            base = (const char*)(g + (int64_t)nz * plane * lda);
        }
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            if (tid + j * A3_THREADS < A3_PL) {
                const ValT* src = (const ValT*)(base + soff[j]);
#pragma unroll
                for (int c = 0; c < LDA; ++c) stage[j][c] = src[c];
            }
        }
    };
    // BOXED: plane n of the ring read through box_cotangent straight into LDS.
    // When every node of the ring plane lies inside the real range on every
    // axis (0 < u < N: interior tiles and planes), each node holds its own
    // cotangent and has no other stored copy: one load per node.  Otherwise
    // (tiles on the periodic faces) every node goes through box_cotangent.
    auto fetch_store_boxed = [&](int n) {
        ValT* dst = ring + (n & 3) * (LDA * A3_PL);
        const bool inner = tx0 - 1 > 0 && tx0 + 14 < m.N[0] && ty0 - 1 > 0 && ty0 + A3_RY - 2 < m.N[1] && n > 0 &&
                           n < m.N[2];
        if (inner) {
            const ValT* src = g + (((int64_t)(n - fg.lo[2]) * fg.ext[1] + (ty0 - 1 - fg.lo[1])) * fg.ext[0] +
                                   (tx0 - 1 - fg.lo[0])) * fg.lda + c0;
            for (int i = tid; i < A3_PL * LDA; i += A3_THREADS) {
                const int node = i / LDA;
                dst[(i % LDA) * A3_PL + node] = src[((int64_t)(node >> 4) * fg.ext[0] + (node & 15)) * fg.lda + i % LDA];
            }
            return;
        }
        for (int i = tid; i < A3_PL * LDA; i += A3_THREADS) {
            const int node = i / LDA;
            const int u[3] = {tx0 - 1 + (node & 15), ty0 - 1 + (node >> 4), n};
            dst[(i % LDA) * A3_PL + node] = (ValT)box_cotangent<ValT>(m, fg, g, c0 + i % LDA, u);
        }
    };
    auto store_plane = [&](int n) {
        ValT* dst = ring + (n & 3) * (LDA * A3_PL);
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            const int node = tid + j * A3_THREADS;
            if (node < A3_PL) {
#pragma unroll
                for (int c = 0; c < LDA; ++c) dst[c * A3_PL + node] = stage[j][c];
            }
        }
    };
    for (int n = z0 - 1; n <= z0 + 2; ++n) {
        if constexpr (BOXED) {
            fetch_store_boxed(n);
        } else {
            fetch_plane(n);
            store_plane(n);
        }
    }

    const bool plane_lane = ex >= 1 && ex <= nxo && row < nyo;
    const int cx = tx0 + ex - 1, cy = ty0 + row;
    const int key0 = (int)((int64_t)cy * Nx + cx);
    // Cell heads run two planes ahead of the computation, the permutation
    // entry of a plane's first particle one plane ahead, and that particle's
    // position and strengths are loaded at the end of the plane before (while
    // the workgroup refills the ring; they are not live during the rounds).
    // (f, n) is the cell's range [f, n) of the permutation, or (TABLE) its
    // key and count word, with t its first particle.
    auto heads = [&](int z, int& f, int& n, int& t) {
        f = 0;
        n = 0;
        t = 0;
        if (plane_lane && z < z1) {
            const int key = (int)((int64_t)z * plane) + key0;
            if constexpr (TABLE) {
                f = key;
                n = (int)bt.count[count_word(key)];   // decoded where used: the load stays in flight
                t = bt.table[key];
            } else {
                f = start[key];
                n = start[key + 1];   // end, subtracted where used: the load stays in flight
            }
        }
    };
    auto count_of = [&](int f, int n) {
        if constexpr (TABLE) return (int)(((unsigned)n >> count_shift(f)) & 255u);
        else return n - f;
    };
    // the cell's first particle, and its particle r < count
    auto lookup = [&](int f, int n, int t) {
        if constexpr (TABLE) return count_of(f, n) > 0 ? t : 0;
        else return n > f ? (perm ? perm[f] : f) : 0;
    };
    auto entry = [&](int f, int r) {
        if constexpr (TABLE) return bt.table[(int64_t)r * bt.ncells + f];
        else return perm ? perm[f + r] : f + r;
    };
    auto particle = [&](int p, int n, PosT (&x)[3], ValT (&u)[LDA]) {
        x[0] = x[1] = x[2] = PosT(0);
#pragma unroll
        for (int c = 0; c < LDA; ++c) u[c] = ValT(0);
        if (n > 0) {
            x[0] = xp[(int64_t)p * 3];
            x[1] = xp[(int64_t)p * 3 + 1];
            x[2] = xp[(int64_t)p * 3 + 2];
#pragma unroll
            for (int c = 0; c < LDA; ++c) u[c] = up[(int64_t)p * lda + c0 + c];
        }
    };
    int first_n, cnt_n, t_n, first_nn, cnt_nn, t_nn;
    heads(z0, first_n, cnt_n, t_n);
    heads(z0 + 1, first_nn, cnt_nn, t_nn);
    int pn = lookup(first_n, cnt_n, t_n);
    PosT xn[3];
    ValT un[LDA];
    particle(pn, count_of(first_n, cnt_n), xn, un);
    __syncthreads();

    for (int cz = z0; cz < z1; ++cz) {
        const int first = first_n, cnt = count_of(first_n, cnt_n);
        // take this plane's heads before the next loads are issued, so that
        // waiting for them does not wait for those loads too
        asm volatile("" ::"v"(first), "v"(cnt) : "memory");
        first_n = first_nn;
        cnt_n = cnt_nn;
        t_n = t_nn;
        const int pnext = lookup(first_n, cnt_n, t_n);   // next plane's first particle
        heads(cz + 2, first_nn, cnt_nn, t_nn);
        const bool more = cz + 1 < z1;
        if (more && staged(cz + 3)) fetch_plane(cz + 3);
        auto round = [&](int p, bool has, const PosT (&x)[3], const ValT (&uv)[LDA]) {
            AccT wx[4] = {0, 0, 0, 0}, wy[4] = {0, 0, 0, 0}, wz[4] = {0, 0, 0, 0};
            AccT dx[4] = {0, 0, 0, 0}, dy[4] = {0, 0, 0, 0}, dz[4] = {0, 0, 0, 0};
            AccT u[LDA];
#pragma unroll
            for (int c = 0; c < LDA; ++c) u[c] = AccT(0);
            if (has) {
                int ci;
                const AccT f0 = (AccT)cell_offset<PosT>(x[0], (PosT)m.minv[0], (PosT)m.dxi[0], ci);
                const AccT f1 = (AccT)cell_offset<PosT>(x[1], (PosT)m.minv[1], (PosT)m.dxi[1], ci);
                const AccT f2 = (AccT)cell_offset<PosT>(x[2], (PosT)m.minv[2], (PosT)m.dxi[2], ci);
                mp4_factored<AccT>(f0, wx);
                mp4_factored<AccT>(f1, wy);
                mp4_factored<AccT>(f2, wz);
                mp4_factored_deriv<AccT>(f0, dx);
                mp4_factored_deriv<AccT>(f1, dy);
                mp4_factored_deriv<AccT>(f2, dz);
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    dx[o] *= (AccT)m.dxi[0];
                    dy[o] *= (AccT)m.dxi[1];
                    dz[o] *= (AccT)m.dxi[2];
                }
#pragma unroll
                for (int c = 0; c < LDA; ++c) u[c] = (AccT)uv[c];
            }
            // One pass per stencil x-offset k: the y and z weights (and
            // derivatives) of the particle k-1 lanes to the right arrive by DPP,
            // every component's column is contracted with them, and the three
            // partial sums go back to that particle at once, where its x weights
            // finish them.  Only one offset's weights are live at a time, so all
            // components fit in one launch (the columns are read once per offset).
            AccT gu[LDA];
#pragma unroll
            for (int c = 0; c < LDA; ++c) gu[c] = AccT(0);
            AccT gx0 = AccT(0), gx1 = AccT(0), gx2 = AccT(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                AccT a[4], da[4], b[4], db[4];
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    a[o] = k == 0 ? row_shl<1>(wy[o]) : k == 1 ? wy[o] : k == 2 ? row_shr<1>(wy[o]) : row_shr<2>(wy[o]);
                    da[o] = k == 0 ? row_shl<1>(dy[o]) : k == 1 ? dy[o] : k == 2 ? row_shr<1>(dy[o]) : row_shr<2>(dy[o]);
                    b[o] = k == 0 ? row_shl<1>(wz[o]) : k == 1 ? wz[o] : k == 2 ? row_shr<1>(wz[o]) : row_shr<2>(wz[o]);
                    db[o] = k == 0 ? row_shl<1>(dz[o]) : k == 1 ? dz[o] : k == 2 ? row_shr<1>(dz[o]) : row_shr<2>(dz[o]);
                }
#pragma unroll
                for (int c = 0; c < LDA; ++c) {
                    AccT C0 = AccT(0), Cy = AccT(0), Cz = AccT(0);
#pragma unroll
                    for (int oz = 0; oz < 4; ++oz) {
                        const ValT* f = ring + ((cz - 1 + oz) & 3) * (LDA * A3_PL) + c * A3_PL + row * 16 + ex;
                        const AccT g0 = (AccT)f[0], g1 = (AccT)f[16], g2 = (AccT)f[32], g3 = (AccT)f[48];
                        AccT t = a[0] * g0;
                        t += a[1] * g1;
                        t += a[2] * g2;
                        t += a[3] * g3;
                        AccT td = da[0] * g0;
                        td += da[1] * g1;
                        td += da[2] * g2;
                        td += da[3] * g3;
                        C0 += b[oz] * t;
                        Cz += db[oz] * t;
                        Cy += b[oz] * td;
                    }
                    // back to the particle: the sums of lane P-1+k
                    const AccT s0 = k == 0 ? row_shr<1>(C0) : k == 1 ? C0 : k == 2 ? row_shl<1>(C0) : row_shl<2>(C0);
                    const AccT sy = k == 0 ? row_shr<1>(Cy) : k == 1 ? Cy : k == 2 ? row_shl<1>(Cy) : row_shl<2>(Cy);
                    const AccT sz = k == 0 ? row_shr<1>(Cz) : k == 1 ? Cz : k == 2 ? row_shl<1>(Cz) : row_shl<2>(Cz);
                    gu[c] += wx[k] * s0;
                    gx0 += u[c] * dx[k] * s0;
                    gx1 += u[c] * wx[k] * sy;
                    gx2 += u[c] * wx[k] * sz;
                    // one component at a time (register pressure)
                    asm volatile("" : "+v"(gx0), "+v"(gx1), "+v"(gx2)::"memory");
                }
            }
            if (has && grad_u) {
#pragma unroll
                for (int c = 0; c < LDA; ++c) grad_u[(int64_t)p * lda + c0 + c] = (ValT)gu[c];
            }
            if (has && grad_x) {
                ValT* o = grad_x + (int64_t)p * 3;
                if (accumulate_x) {
                    o[0] = (ValT)((AccT)o[0] + gx0);
                    o[1] = (ValT)((AccT)o[1] + gx1);
                    o[2] = (ValT)((AccT)o[2] + gx2);
                } else {
                    o[0] = (ValT)gx0;
                    o[1] = (ValT)gx1;
                    o[2] = (ValT)gx2;
                }
            }
        };
        if (__ballot(cnt > 0) != 0ull) round(pn, cnt > 0, xn, un);
        for (int r = 1; __ballot(cnt > r) != 0ull; ++r) {
            PosT x[3];
            ValT u[LDA];
            const int p = cnt > r ? entry(first, r) : 0;
            particle(p, cnt - r, x, u);
            round(p, cnt > r, x, u);
        }
        pn = pnext;
        particle(pn, count_of(first_n, cnt_n), xn, un);   // next plane's round 0
        __syncthreads();   // every row is done with plane cz-1: its slot takes plane cz+3
        if (more) {
            if (staged(cz + 3)) store_plane(cz + 3);
            else fetch_store_boxed(cz + 3);
        }
        __syncthreads();
    }
}

template <int LL, typename PosT, typename ValT, typename AccT, bool BOXED, bool TABLE = false>
static void launch_a3_form(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const ValT* g,
                           const FieldGeom& fg, ValT* gu, ValT* gx, int lda, int c0, int acc, hipStream_t s) {
    static std::atomic<int> slots{0};
    if (slots.load() == 0)
        slots.store(resident_workgroups(k_a3_tile<LL, PosT, ValT, AccT, BOXED, TABLE>, A3_THREADS));
    const int tx = (m.N[0] + A3_TX - 1) / A3_TX, ty = (m.N[1] + A3_TY - 1) / A3_TY;
    const int tz = pick_z_chunk_balanced((int64_t)tx * ty, m.N[2], slots.load(), 32);
    dim3 grid((unsigned)((int64_t)tx * ty * ((m.N[2] + tz - 1) / tz)));
    const BinTable bt{plan->count.as<uint32_t>(), plan->table.as<int>(), plan->ncells};
    hipLaunchKernelGGL((k_a3_tile<LL, PosT, ValT, AccT, BOXED, TABLE>), grid, dim3(A3_THREADS), 0, s, m,
                       plan->start.as<int>(), tile_perm(plan), bt, xp, up, g, fg, gu, gx, lda, c0, tz, acc);
}

// The compact layout of every axis, or a stored box on some axis.
static bool a3_compact(const DevMesh& m, const FieldGeom& fg) {
    for (int d = 0; d < 3; ++d)
        if (fg.lo[d] != 0 || fg.ext[d] != m.N[d]) return false;
    return true;
}

template <int LL, typename PosT, typename ValT, typename AccT>
static void launch_a3(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const ValT* g,
                      const FieldGeom& fg, ValT* gu, ValT* gx, int lda, int c0, int acc, hipStream_t s) {
    const bool compact = a3_compact(m, fg);
    if (plan->bin_table) {
        if (compact) launch_a3_form<LL, PosT, ValT, AccT, false, true>(plan, m, xp, up, g, fg, gu, gx, lda, c0, acc, s);
        else launch_a3_form<LL, PosT, ValT, AccT, true, true>(plan, m, xp, up, g, fg, gu, gx, lda, c0, acc, s);
    } else {
        if (compact) launch_a3_form<LL, PosT, ValT, AccT, false>(plan, m, xp, up, g, fg, gu, gx, lda, c0, acc, s);
        else launch_a3_form<LL, PosT, ValT, AccT, true>(plan, m, xp, up, g, fg, gu, gx, lda, c0, acc, s);
    }
}

// One launch per three components (the position gradient accumulates over
// launches when lda > 3: four components in one launch would spill); inside a
// launch the components are contracted one after another and the stencil
// x-offsets one per pass, which keeps the kernel at 3 wavefronts per SIMD.
template <typename PosT, typename ValT>
static int adjoint3d_typed(ppmg_plan* plan, const DevMesh& m, int64_t np, const PosT* xp, int lda, const ValT* up,
                           const ValT* g, const FieldGeom& fg, ValT* gu, ValT* gx, hipStream_t s) {
    (void)np;
    for (int c0 = 0; c0 < lda; c0 += 3) {
        const int L = lda - c0 < 3 ? lda - c0 : 3;
        const int acc = c0 > 0 ? 1 : 0;
#define AL(LL) launch_a3<LL, PosT, ValT, double>(plan, m, xp, up, g, fg, gu, gx, lda, c0, acc, s)
        if (L == 1) AL(1);
        else if (L == 2) AL(2);
        else AL(3);
#undef AL
        PPMG_HIP(hipGetLastError());
    }
    return 0;
}

// Periodic axes, each in the compact layout or a stored box holding the real
// range 0..N (the reference's 1-g..Nm+g box is lo = -g).
bool adjoint3d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg) {
    if (m.dim != 3 || kernel != PPMG_KERNEL_MP4) return false;
    for (int d = 0; d < 3; ++d) {
        if (!m.periodic[d] || m.N[d] < 4) return false;
        const bool compact = fg.lo[d] == 0 && fg.ext[d] == m.N[d];
        const bool boxed = fg.lo[d] <= 0 && fg.lo[d] + fg.ext[d] - 1 >= m.N[d];
        if (!compact && !boxed) return false;
    }
    // the compact ring fetch addresses a field plane with 32-bit byte offsets
    return !a3_compact(m, fg) || (int64_t)m.N[0] * m.N[1] * fg.lda * 8 < ((int64_t)1 << 31);
}

// Uses the forward's binning in the plan (every particle is inside the sub:
// the forward P2M would have failed otherwise).
int adjoint3d_fast(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp, int lda, const void* up,
                   const FieldGeom& fg, const void* g, void* gu, void* gx, hipStream_t s) {
    if (np <= 0) return 0;
    if (prec == PPMG_PREC_F64)
        return adjoint3d_typed<double, double>(plan, m, np, (const double*)xp, lda, (const double*)up,
                                               (const double*)g, fg, (double*)gu, (double*)gx, s);
    if (prec == PPMG_PREC_F32C)
        return adjoint3d_typed<double, float>(plan, m, np, (const double*)xp, lda, (const float*)up,
                                              (const float*)g, fg, (float*)gu, (float*)gx, s);
    return adjoint3d_typed<float, float>(plan, m, np, (const float*)xp, lda, (const float*)up, (const float*)g, fg,
                                         (float*)gu, (float*)gx, s);
}

}  // namespace ppmg
