// ppmg_p2m2d.hip -- 2D P2M fast path for gfx950 (M'4 and BSP2): tiled
// y-march deposit.
//
// Replaces the 2D scatter loops of ppm_interp_p2m (ppm_interp_p2m.f:3574-3663
// for M'4, :3057-3104 for BSP2's weights) plus the periodic ghost fold
// (:3670-3716), on periodic and open axes, written in the compact layout or
// into a stored box holding the real range (field_up(1-g:Nm+g)).
//
// Work decomposition (one wavefront per workgroup):
//   * a wavefront owns a strip of TT = 16 - H node columns (13 for M'4, 15 for
//     BSP2; H = KS - 1 halo cells) and a chunk of node rows [y0, y1);
//   * lane ex of DPP row erow stands for cell column tx0 - (H - B) + ex; a step
//     takes the four cell rows 4s + erow of the strip's window, one per DPP
//     row, and every lane walks its cell's particles one per round;
//   * lane ex finishes node column ex - H from the particles of lanes ex-k
//     (k = 0..H, fetched by DPP row shifts), and adds one value per (y offset,
//     component) to an LDS ring of 8 node rows;
//   * after step s the node rows 4s - H .. 4s - H + 3 of the chunk have all
//     their cells and are written out with coalesced row stores.
// Lanes of one instruction address distinct nodes and one wavefront owns the
// ring, so the order of every node's additions is a fixed function of the
// binning: results are bitwise reproducible, with no global float atomics.
#include <atomic>

#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"
#include "ppmg_tilebox.h"

namespace ppmg {

constexpr int P2_SERIAL = 4;   // later rounds with at most this many active lanes go one particle at a time
constexpr int P2_RING = 8;     // node rows in the ring (a step touches 4 + H <= 7)

// TABLE: the binning is a slot table (byte counts, particle r of cell c at
// bt.table[r * ncells + c]) instead of heads and a permutation; a cell's
// particles come in the same order.
template <int KS, int LDA, typename PosT, typename ValT, bool OPEN, bool TABLE>
__global__ void __launch_bounds__(64)
k_p2_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, BinTable bt,
          const PosT* __restrict__ xp, const ValT* __restrict__ up, FieldGeom fg, int c0, int ty,
          ValT* __restrict__ field) {
    constexpr int H = Stencil<KS>::H, B = Stencil<KS>::B;
    constexpr int TT = 16 - H;                    // owned node columns per strip
    constexpr int ROW = TT * LDA;                 // ring slots per node row: [c][x]
    static_assert(ROW <= 64, "a node row is written by one pass of the wavefront");
    __shared__ double ring[P2_RING * ROW];
    const int lane = threadIdx.x;
    const int lda = fg.lda;
    // nodes the tiles produce per axis: the N unique nodes of a periodic axis,
    // the N+1 real nodes 0..N of an open one
    const int Rx = m.N[0] + !m.periodic[0], Ry = m.N[1] + !m.periodic[1];
    const int ntx = (Rx + TT - 1) / TT;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int tx0 = (int)(tile % ntx) * TT;
    const int y0 = (int)(tile / ntx) * ty;
    const int y1 = min(y0 + ty, Ry);
    const int nxo = min(TT, Rx - tx0), nyo = y1 - y0;
    for (int i = lane; i < P2_RING * ROW; i += 64) ring[i] = 0.0;

    const int ex = lane & 15, erow = lane >> 4;
    const int lx = ex - H;                          // node column this lane owns after the row sum
    const bool xown = lx >= 0 && lx < nxo;
    auto wcell = [&](int c, int d) {
        const int nb = m.nbin[d];
        if (!OPEN || m.periodic[d]) {
            c %= nb;
            return c < 0 ? c + nb : c;
        }
        return c >= 0 && c < nb ? c : -1;
    };
    const int cx = wcell(tx0 - (H - B) + ex, 0);
    // cell rows of the window: y0 - (H - B) .. y1 - 1 + B, four per step
    const int nsteps = (nyo + H + 3) / 4;
    // (f, n): the cell's range [f, n) of the permutation, or (TABLE) its key
    // and count word, with t its first particle
    auto heads = [&](int st, int& f, int& n, int& t) {
        f = 0;
        n = 0;
        t = 0;
        const int r = 4 * st + erow;                // window row
        const int cr = y0 - (H - B) + r;
        if (st < nsteps && cr <= y1 - 1 + B) {
            const int cy = wcell(cr, 1);
            if (!OPEN || (cx >= 0 && cy >= 0)) {
                const int key = cx + m.nbin[0] * cy;
                if constexpr (TABLE) {
                    f = key;
                    n = (int)bt.count[count_word(key)];   // decoded where used: the load stays in flight
                    t = bt.table[key];
                } else {
                    f = start[key];
                    n = start[key + 1];   // end, subtracted where used: the load stays in flight
                }
            }
        }
    };
    auto count_of = [&](int f, int n) {
        if constexpr (TABLE) return (int)(((unsigned)n >> count_shift(f)) & 255u);
        else return n - f;
    };
    auto entry = [&](int f, int r) {
        if constexpr (TABLE) return bt.table[(int64_t)r * bt.ncells + f];
        else return perm ? perm[f + r] : f + r;
    };
    auto load = [&](int p, bool has, PosT (&x)[2], ValT (&u)[LDA]) {
        x[0] = x[1] = PosT(0);
#pragma unroll
        for (int c = 0; c < LDA; ++c) u[c] = ValT(0);
        if (has) {
            x[0] = xp[(int64_t)p * 2];
            x[1] = xp[(int64_t)p * 2 + 1];
#pragma unroll
            for (int c = 0; c < LDA; ++c) u[c] = up[(int64_t)p * lda + c0 + c];
        }
    };
    auto offsets = [&](const PosT (&x)[2], ValT (&fr)[2]) {
        int ci;
        fr[0] = (ValT)cell_offset<PosT>(x[0], (PosT)m.minv[0], (PosT)m.dxi[0], ci);
        fr[1] = (ValT)cell_offset<PosT>(x[1], (PosT)m.minv[1], (PosT)m.dxi[1], ci);
    };
    auto lookup = [&](int f, int n, int t) {
        if constexpr (TABLE) return count_of(f, n) > 0 ? t : 0;
        else return n > f ? (perm ? perm[f] : f) : 0;
    };
    int f1, n1, t1, f2, n2, t2;
    heads(0, f1, n1, t1);
    heads(1, f2, n2, t2);
    int pn = lookup(f1, n1, t1);
    PosT xn[2];
    ValT un[LDA];
    load(pn, count_of(f1, n1) > 0, xn, un);

    for (int st = 0; st < nsteps; ++st) {
        const int first = f1, cnt = count_of(f1, n1);
        asm volatile("" ::"v"(first), "v"(cnt) : "memory");
        f1 = f2;
        n1 = n2;
        t1 = t2;
        pn = lookup(f1, n1, t1);
        heads(st + 2, f2, n2, t2);
        const int ly0 = 4 * st + erow - H;          // chunk node row of y offset 0
        // one round: every lane deposits one particle of its cell (zero
        // strength where it has none)
        auto round = [&](const PosT (&x)[2], const ValT (&u)[LDA]) {
            ValT fr[2], wx[KS], wy[KS];
            offsets(x, fr);
            tile_weights<KS, ValT>(fr[0], wx);
            tile_weights<KS, ValT>(fr[1], wy);
            // G[k][c] = wx_k u_c and the y weights of the particle k lanes below
            ValT G[KS][LDA], WY[KS][KS];
#pragma unroll
            for (int c = 0; c < LDA; ++c) {
                G[0][c] = wx[0] * u[c];
                G[1][c] = row_shr<1>(wx[1] * u[c]);
                if constexpr (KS == 4) {
                    G[2][c] = row_shr<2>(wx[2] * u[c]);
                    G[3][c] = row_shr<3>(wx[3] * u[c]);
                }
            }
#pragma unroll
            for (int o = 0; o < KS; ++o) {
                WY[0][o] = wy[o];
                WY[1][o] = row_shr<1>(wy[o]);
                if constexpr (KS == 4) {
                    WY[2][o] = row_shr<2>(wy[o]);
                    WY[3][o] = row_shr<3>(wy[o]);
                }
            }
#pragma unroll
            for (int oy = 0; oy < KS; ++oy) {
                const int ly = ly0 + oy;
                const bool own = xown && ly >= 0 && ly < nyo;
                double* row = ring + (ly & (P2_RING - 1)) * ROW;
#pragma unroll
                for (int c = 0; c < LDA; ++c) {
                    ValT v = WY[0][oy] * G[0][c];
#pragma unroll
                    for (int k = 1; k < KS; ++k) v += WY[k][oy] * G[k][c];
                    if (own) atomicAdd(row + c * TT + lx, (double)v);
                }
            }
        };
        if (__ballot(cnt > 0) != 0ull) round(xn, un);
        // later rounds: as the 3D deposit, one particle at a time once at most
        // P2_SERIAL lanes still have one (lanes 0..KS^2-1 take its nodes)
        for (int r = 1;; ++r) {
            uint64_t act = __ballot(cnt > r);
            if (act == 0ull) break;
            if (__popcll(act) > P2_SERIAL) {
                PosT x[2];
                ValT u[LDA];
                load(cnt > r ? entry(first, r) : 0, cnt > r, x, u);
                round(x, u);
                continue;
            }
            const int ox = lane & (KS - 1), oy = (lane / KS) & (KS - 1);
            const bool slane = lane < KS * KS;
            do {
                const int src = __builtin_ctzll(act);
                act &= act - 1;
                const int sf = __builtin_amdgcn_readlane(first, src);
                const int sc = __builtin_amdgcn_readlane(cnt, src);
                const int nx = (src & 15) - H + ox, ny = 4 * st + (src >> 4) - H + oy;
                const bool own = slane && nx >= 0 && nx < nxo && ny >= 0 && ny < nyo;
                double* row = ring + (ny & (P2_RING - 1)) * ROW;
                for (int rr = r; rr < sc; ++rr) {
                    PosT x[2];
                    ValT u[LDA], fr[2];
                    load(entry(sf, rr), true, x, u);
                    offsets(x, fr);
                    const ValT w = tile_weight_at<KS, ValT>(fr[0], ox) * tile_weight_at<KS, ValT>(fr[1], oy);
                    if (own) {
#pragma unroll
                        for (int c = 0; c < LDA; ++c) atomicAdd(row + c * TT + nx, (double)(w * u[c]));
                    }
                }
            } while (act != 0ull);
            break;
        }
        load(pn, count_of(f1, n1) > 0, xn, un);
        // node rows 4s-H .. 4s-H+3 have all their cells: write the owned ones
        // and clear their slots for rows 4 steps on
        wave_lds_sync();
        const int wx = lane / LDA, wc = lane % LDA;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ly = 4 * st - H + q;
            double* row = ring + (ly & (P2_RING - 1)) * ROW;
            if (lane < ROW) {
                const double v = row[wc * TT + wx];
                row[wc * TT + wx] = 0.0;
                if (ly >= 0 && ly < nyo && wx < nxo) {
                    const int64_t at = ((int64_t)(y0 + ly - fg.lo[1]) * fg.ext[0] + (tx0 + wx - fg.lo[0])) * lda;
                    field[at + c0 + wc] = (ValT)v;
                }
            }
        }
        wave_lds_sync();
    }
}

// The 2D deposit (and its stored box's shell) reads a slot-table binning on
// a periodic mesh.
bool p2m2d_table_applicable(const DevMesh& m, int kernel, const FieldGeom& fg) {
    return p2m2d_fast_applicable(m, kernel, fg) && m.periodic[0] && m.periodic[1];
}

// Per axis the compact layout of a periodic axis (the N unique nodes) or a
// stored box holding the whole real range 0..N.
bool p2m2d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg) {
    if (m.dim != 2 || (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)) return false;
    for (int d = 0; d < 2; ++d) {
        if (m.N[d] < 4) return false;
        const bool compact = m.periodic[d] && fg.lo[d] == 0 && fg.ext[d] == m.N[d];
        const bool boxed = fg.lo[d] <= 0 && fg.lo[d] + fg.ext[d] - 1 >= m.N[d];
        if (!compact && !boxed) return false;
    }
    return true;
}

// One launch for components c0 .. c0+LL-1.  The chunk height comes from the
// resident workgroup count (each chunk marches H extra cell rows).
template <int KS, int LL, typename PosT, typename ValT, bool OPEN, bool TABLE = false>
static void launch_p2_open(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                           const FieldGeom& fg, int c0, ValT* field, hipStream_t s) {
    constexpr int H = Stencil<KS>::H, TT = 16 - H;
    static std::atomic<int> slots{0};
    if (slots.load() == 0) slots.store(resident_workgroups(k_p2_tile<KS, LL, PosT, ValT, OPEN, TABLE>, 64));
    const int Rx = m.N[0] + !m.periodic[0], Ry = m.N[1] + !m.periodic[1];
    const int tx = (Rx + TT - 1) / TT;
    const int ty = pick_z_chunk(tx, Ry, slots.load(), H);
    dim3 grid((unsigned)((int64_t)tx * ((Ry + ty - 1) / ty)));
    const BinTable bt{plan->count.as<uint32_t>(), plan->table.as<int>(), plan->ncells};
    hipLaunchKernelGGL((k_p2_tile<KS, LL, PosT, ValT, OPEN, TABLE>), grid, dim3(64), 0, s, m, plan->start.as<int>(),
                       perm, bt, xp, up, fg, c0, ty, field);
}

template <int KS, typename PosT, typename ValT>
static int p2m2d_typed(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const FieldGeom& fg, const ValT* up,
                       ValT* field, hipStream_t s) {
    mark(plan, s);
    const int lda = fg.lda;
    const bool open = !(m.periodic[0] && m.periodic[1]);
    // a slot table is binned on periodic meshes only (p2m2d_table_applicable)
    if (plan->bin_table && open && materialize_heads(plan, s)) return 1;
    const int* pk = tile_perm(plan);
    for (int c0 = 0; c0 < lda; c0 += 4) {
        const int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL)                                                                              \
    do {                                                                                    \
        if (plan->bin_table) launch_p2_open<KS, LL, PosT, ValT, false, true>(plan, m, xp, up, pk, fg, c0, field, s); \
        else if (open) launch_p2_open<KS, LL, PosT, ValT, true>(plan, m, xp, up, pk, fg, c0, field, s); \
        else launch_p2_open<KS, LL, PosT, ValT, false>(plan, m, xp, up, pk, fg, c0, field, s);     \
    } while (0)
        if (L == 1) TL(1);
        else if (L == 2) TL(2);
        else if (L == 3) TL(3);
        else TL(4);
#undef TL
        PPMG_HIP(hipGetLastError());
    }
    const WTab<ValT> wt{{nullptr, nullptr, nullptr}};
    if (fill_stored_box<2, KS, PosT, ValT>(plan, m, xp, up, fg, field, wt, s)) return 1;
    mark(plan, s);
    return 0;
}

int p2m2d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, const void* xp, const FieldGeom& fg,
               const void* up, void* field, hipStream_t s) {
#define TYPED(KS)                                                                                           \
    do {                                                                                                    \
        if (prec == PPMG_PREC_F64)                                                                          \
            return p2m2d_typed<KS, double, double>(plan, m, (const double*)xp, fg, (const double*)up,      \
                                                   (double*)field, s);                                      \
        if (prec == PPMG_PREC_F32C)                                                                         \
            return p2m2d_typed<KS, double, float>(plan, m, (const double*)xp, fg, (const float*)up,        \
                                                  (float*)field, s);                                        \
        return p2m2d_typed<KS, float, float>(plan, m, (const float*)xp, fg, (const float*)up, (float*)field, s); \
    } while (0)
    if (kernel == PPMG_KERNEL_BSP2) TYPED(2);
    TYPED(4);
#undef TYPED
}

}  // namespace ppmg
