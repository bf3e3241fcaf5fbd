// ppmg_m2p2d.hip -- 2D M2P fast path for gfx950 (M'4 and BSP2): tiled
// y-march gather with DPP row exchange.
//
// Replaces the 2D gather loops of ppm_interp_m2p (ppm_interp_m2p.f:809-890 for
// M'4, :3245-3528 for BSP2) on periodic axes (compact layout or a stored box
// read with ghost_get's node map) and open axes (a stored box holding every
// node a stencil reaches).
//
// A wavefront (one per workgroup) owns the particles of a strip of TX = 17 -
// KS cell columns (13 for M'4, 15 for BSP2) over a chunk of cell rows, taken
// four rows a step (one per 16-lane DPP row).  Lane ex stands for node column
// tx0 - B + ex and cell column tx0 - B + ex (lanes B .. B+TX-1 carry
// particles).  The node rows the steps read are staged in an LDS ring of 8
// rows (component-major rows of 16 nodes); after a step its four oldest rows
// are replaced by the four the step after next needs, loaded into registers
// while the step computes.  As in the 3D gather, every lane contracts ONE
// column of its node column with the y weights of the KS particles whose
// x-stencil covers it (fetched by DPP), and the partial sums go back by the
// opposite shifts, where the x weights finish each particle's sum.
#include <atomic>

#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

constexpr int M2_RING = 8;   // node rows in the ring

// TABLE: the binning is a slot table (byte counts, particle r of cell c at
// bt.table[r * ncells + c]) instead of heads and a permutation.
template <int KS, int LDA, typename PosT, typename ValT, typename AccT, bool OPEN, bool TABLE>
__global__ void __launch_bounds__(64)
k_m2_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, BinTable bt,
          const PosT* __restrict__ xp, const ValT* __restrict__ field, FieldGeom fg, ValT* __restrict__ up, int lda,
          int c0, int ty) {
    constexpr int B = Stencil<KS>::B, TX = 17 - KS;
    constexpr int ROWV = 16 * LDA;                 // values per ring row: [c][nx]
    __shared__ ValT ring[M2_RING * ROWV];
    const int lane = threadIdx.x;
    const int ex = lane & 15, erow = lane >> 4;
    const int Nx = m.nbin[0], Ny = m.nbin[1];      // cells: N per periodic axis, N+1 per open one
    const int ntx = (Nx + TX - 1) / TX;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int tx0 = (int)(tile % ntx) * TX;
    const int y0 = (int)(tile / ntx) * ty;
    const int y1 = min(y0 + ty, Ny);
    const int nxo = min(TX, Nx - tx0);
    const int nsteps = (y1 - y0 + 3) / 4;

    // ring row r holds node row y0 - B + r; lane (q, nx) stages node column
    // nx of row r = 4j + q, all LDA components (adjacent in the field)
    const int scol = stored_node<OPEN>(m, fg, 0, tx0 - B + ex) - fg.lo[0];
    auto fetch_row = [&](int r, ValT (&v)[LDA]) {
        const int srow = stored_node<OPEN>(m, fg, 1, y0 - B + r) - fg.lo[1];
        const ValT* src = field + ((int64_t)srow * fg.ext[0] + scol) * lda + c0;
#pragma unroll
        for (int c = 0; c < LDA; ++c) v[c] = src[c];
    };
    auto store_row = [&](int r, const ValT (&v)[LDA]) {
        ValT* dst = ring + (r & (M2_RING - 1)) * ROWV + ex;
#pragma unroll
        for (int c = 0; c < LDA; ++c) dst[c * 16] = v[c];
    };
    {
        ValT v[LDA], w[LDA];
        fetch_row(erow, v);
        fetch_row(4 + erow, w);
        store_row(erow, v);
        store_row(4 + erow, w);
    }

    const bool plane_lane = ex >= B && ex - B < nxo;   // lanes that carry particles (row checked per step)
    const int cx = tx0 - B + ex;
    // (f, n): the cell's range [f, n) of the permutation, or (TABLE) its key
    // and count word, with t its first particle
    auto heads = [&](int st, int& f, int& n, int& t) {
        f = 0;
        n = 0;
        t = 0;
        const int cy = y0 + 4 * st + erow;
        if (plane_lane && st < nsteps && cy < y1) {
            const int key = cx + Nx * cy;
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
    auto lookup = [&](int f, int n, int t) {
        if constexpr (TABLE) return count_of(f, n) > 0 ? t : 0;
        else return n > f ? (perm ? perm[f] : f) : 0;
    };
    auto entry = [&](int f, int r) {
        if constexpr (TABLE) return bt.table[(int64_t)r * bt.ncells + f];
        else return perm ? perm[f + r] : f + r;
    };
    auto position = [&](int p, int n, PosT (&x)[2]) {
        x[0] = x[1] = PosT(0);
        if (n > 0) {
            x[0] = xp[(int64_t)p * 2];
            x[1] = xp[(int64_t)p * 2 + 1];
        }
    };
    int first_n, cnt_n, t_n, first_nn, cnt_nn, t_nn;
    heads(0, first_n, cnt_n, t_n);
    heads(1, first_nn, cnt_nn, t_nn);
    int pn = lookup(first_n, cnt_n, t_n);
    PosT xn[2];
    position(pn, count_of(first_n, cnt_n), xn);
    wave_lds_sync();

    for (int st = 0; st < nsteps; ++st) {
        const int first = first_n, cnt = count_of(first_n, cnt_n);
        asm volatile("" ::"v"(first), "v"(cnt) : "memory");
        first_n = first_nn;
        cnt_n = cnt_nn;
        t_n = t_nn;
        const int pnext = lookup(first_n, cnt_n, t_n);
        heads(st + 2, first_nn, cnt_nn, t_nn);
        // the four node rows the step after this one adds to the ring
        const bool more = st + 1 < nsteps;
        ValT stage[LDA];
        if (more) fetch_row(4 * st + 8 + erow, stage);
        auto round = [&](int p, bool has, const PosT (&x)[2]) {
            AccT wx[KS], wy[KS];
#pragma unroll
            for (int o = 0; o < KS; ++o) wx[o] = wy[o] = AccT(0);
            if (has) {
                int ci;
                tile_weights<KS, AccT>((AccT)cell_offset<PosT>(x[0], (PosT)m.minv[0], (PosT)m.dxi[0], ci), wx);
                tile_weights<KS, AccT>((AccT)cell_offset<PosT>(x[1], (PosT)m.minv[1], (PosT)m.dxi[1], ci), wy);
            }
            // y weights of the particle whose stencil takes this lane's column as
            // x offset k: that of lane ex + B - k
            AccT a[KS][KS];
#pragma unroll
            for (int o = 0; o < KS; ++o) {
                a[0][o] = lane_from<B>(wy[o]);
                a[1][o] = lane_from<B - 1>(wy[o]);
                if constexpr (KS == 4) {
                    a[2][o] = lane_from<B - 2>(wy[o]);
                    a[3][o] = lane_from<B - 3>(wy[o]);
                }
            }
            ValT res[LDA];
#pragma unroll
            for (int c = 0; c < LDA; ++c) {
                AccT fv[KS];
#pragma unroll
                for (int oy = 0; oy < KS; ++oy)
                    fv[oy] = (AccT)ring[((4 * st + erow + oy) & (M2_RING - 1)) * ROWV + c * 16 + ex];
                AccT C[KS];
#pragma unroll
                for (int k = 0; k < KS; ++k) {
                    C[k] = a[k][0] * fv[0];
#pragma unroll
                    for (int oy = 1; oy < KS; ++oy) C[k] += a[k][oy] * fv[oy];
                }
                // x offset k of this lane's particle is the column of lane ex - B + k
                AccT u = wx[0] * lane_from<-B>(C[0]);
                u += wx[1] * lane_from<1 - B>(C[1]);
                if constexpr (KS == 4) {
                    u += wx[2] * lane_from<2 - B>(C[2]);
                    u += wx[3] * lane_from<3 - B>(C[3]);
                }
                res[c] = (ValT)u;
            }
            if (has) {
#pragma unroll
                for (int c = 0; c < LDA; ++c) up[(int64_t)p * lda + c0 + c] = res[c];
            }
        };
        if (__ballot(cnt > 0) != 0ull) round(pn, cnt > 0, xn);
        for (int r = 1; __ballot(cnt > r) != 0ull; ++r) {
            PosT x[2];
            const int p = cnt > r ? entry(first, r) : 0;
            position(p, cnt - r, x);
            round(p, cnt > r, x);
        }
        pn = pnext;
        position(pn, count_of(first_n, cnt_n), xn);   // next step's round 0
        // rows 4st .. 4st+3 are done with: their slots take rows 4st+8 .. 4st+11
        wave_lds_sync();
        if (more) store_row(4 * st + 8 + erow, stage);
        wave_lds_sync();
    }
}

// Per periodic axis the compact layout or a stored box holding the whole real
// range (read with ghost_get's node map); per open axis a box holding every
// node a stencil reaches (-1..N+2 for M'4, 0..N+1 for BSP2).  A periodic box
// beside an open axis (where the node map depends on both coordinates) takes
// the general path.
bool m2p2d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg) {
    if (m.dim != 2 || (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)) return false;
    const int below = kernel == PPMG_KERNEL_MP4 ? 1 : 0, above = kernel == PPMG_KERNEL_MP4 ? 2 : 1;
    const bool open = !m.periodic[0] || !m.periodic[1];
    for (int d = 0; d < 2; ++d) {
        if (m.N[d] < 4) return false;
        const int hi = fg.lo[d] + fg.ext[d] - 1;
        if (m.periodic[d]) {
            if (open && fg.swap[d]) return false;
            const bool compact = fg.lo[d] == 0 && fg.ext[d] == m.N[d];
            const bool boxed = fg.lo[d] <= 0 && hi >= m.N[d];
            if (!compact && !boxed) return false;
        } else if (fg.lo[d] > -below || hi < m.N[d] + above) {
            return false;
        }
    }
    return true;
}

template <int KS, int LL, typename PosT, typename ValT, typename AccT, bool OPEN, bool TABLE = false>
static void launch_m2_open(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const FieldGeom& fg, int c0,
                           const ValT* field, ValT* up, hipStream_t s) {
    constexpr int TX = 17 - KS;
    static std::atomic<int> slots{0};
    if (slots.load() == 0) slots.store(resident_workgroups(k_m2_tile<KS, LL, PosT, ValT, AccT, OPEN, TABLE>, 64));
    const int tx = (m.nbin[0] + TX - 1) / TX;
    const int ty = pick_z_chunk(tx, m.nbin[1], slots.load(), 0);
    dim3 grid((unsigned)((int64_t)tx * ((m.nbin[1] + ty - 1) / ty)));
    const BinTable bt{plan->count.as<uint32_t>(), plan->table.as<int>(), plan->ncells};
    hipLaunchKernelGGL((k_m2_tile<KS, LL, PosT, ValT, AccT, OPEN, TABLE>), grid, dim3(64), 0, s, m,
                       plan->start.as<int>(), tile_perm(plan), bt, xp, field, fg, up, fg.lda, c0, ty);
}

template <int KS, typename PosT, typename ValT, typename AccT>
static int m2p2d_typed(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const FieldGeom& fg, const ValT* field,
                       ValT* up, hipStream_t s) {
    const int lda = fg.lda;
    const bool open = !(m.periodic[0] && m.periodic[1]);
    for (int c0 = 0; c0 < lda; c0 += 4) {
        const int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL)                                                                                   \
    do {                                                                                         \
        if (plan->bin_table) launch_m2_open<KS, LL, PosT, ValT, AccT, false, true>(plan, m, xp, fg, c0, field, up, s); \
        else if (open) launch_m2_open<KS, LL, PosT, ValT, AccT, true>(plan, m, xp, fg, c0, field, up, s);  \
        else launch_m2_open<KS, LL, PosT, ValT, AccT, false>(plan, m, xp, fg, c0, field, up, s);      \
    } while (0)
        if (L == 1) TL(1);
        else if (L == 2) TL(2);
        else if (L == 3) TL(3);
        else TL(4);
#undef TL
        PPMG_HIP(hipGetLastError());
    }
    return 0;
}

// Bins the particles into the plan, then gathers.  Returns 1 (nothing done)
// when some particle lies outside the sub, for the general path.  Sums in
// FP64 in every form, as the 3D gather.
int m2p2d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp,
               const FieldGeom& fg, const void* field, void* up, hipStream_t s) {
    int64_t nbad = 0;
    // a slot table on a periodic mesh
    if (bin_particles(plan, m, prec, np, xp, s, &nbad, m.periodic[0] && m.periodic[1])) return PPMG_ERROR_FATAL;
    if (nbad > 0) return 1;
    mark(plan, s);
    if (np <= 0) return 0;
#define TYPED(KS)                                                                                             \
    do {                                                                                                      \
        if (prec == PPMG_PREC_F64)                                                                            \
            return m2p2d_typed<KS, double, double, double>(plan, m, (const double*)xp, fg,                    \
                                                           (const double*)field, (double*)up, s);             \
        if (prec == PPMG_PREC_F32C)                                                                           \
            return m2p2d_typed<KS, double, float, double>(plan, m, (const double*)xp, fg,                     \
                                                          (const float*)field, (float*)up, s);                \
        return m2p2d_typed<KS, float, float, double>(plan, m, (const float*)xp, fg, (const float*)field,      \
                                                     (float*)up, s);                                          \
    } while (0)
    if (kernel == PPMG_KERNEL_BSP2) TYPED(2);
    TYPED(4);
#undef TYPED
}

}  // namespace ppmg
