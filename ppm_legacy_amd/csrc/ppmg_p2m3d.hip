// ppmg_p2m3d.hip -- 3D M'4 P2M fast path for gfx950: tiled z-march deposit.
//
// Replaces the scatter loop of ppm_interp_p2m (ppm_interp_p2m.f:783-1377,
// 2540-3055) plus its periodic ghost fold (:3670-3716) for 3D M'4 on periodic
// and open axes, written in the compact layout (the unique nodes of periodic
// axes) or into a stored box holding the real range (the reference's
// field_up(1-g:Nm+g)), whose node N and ghost layers two small kernels fill.
//
// Work decomposition (three wavefronts per workgroup):
//   * a workgroup owns a 13 x 21 node tile in x,y and a chunk of node planes in z;
//   * it marches the cell planes cz = z0-2 .. z1 (the cells whose M'4 stencil
//     touches its planes) and keeps a ring of 4 node planes in LDS: cell plane
//     cz deposits into node planes cz-1 .. cz+2, after which node plane cz-1
//     is complete and is written out with coalesced row stores;
//   * the 16 x 24 cells whose stencils touch the tile (13 + 3 and 21 + 3 halo)
//     are taken 64 at a time, one cell per lane; each lane walks its cell's
//     particles (binned, in stable order) and adds the owned part of the
//     4x4x4 stencil.
// Lanes of one instruction hold distinct cells, hence write distinct nodes;
// the wavefronts of a workgroup work on disjoint node rows between barriers,
// so the order of every node's additions is a fixed function of the binning,
// and results are bitwise reproducible with no global float atomics.  The
// periodic fold happens through the index wrap of the cells around the tile.
#include <atomic>
#include <cstdlib>

#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"
#include "ppmg_tilebox.h"

namespace ppmg {

constexpr int P3_TT = 13;   // owned nodes per tile along x (16-lane DPP row = 13 + 3 halo cells)
constexpr int P3_TY = 21;   // owned node rows per tile along y (24 cell rows: 6 chunks, 3 wavefronts)
constexpr int P3_SERIAL = 4; // later rounds with at most this many active lanes go one particle at a time

// LDS accumulator of one node component.  Every form accumulates in an FP64
// slot with ds_add_f64 (fire-and-forget, no read-modify-write latency chain):
// the FP64 form adds its FP64 products; the single-precision forms add their
// FP32 products, so each node sum is carried at double width (the same 8 bytes
// a Kahan (sum, carry) float pair occupies, and at least as accurate) and is
// rounded to FP32 once on output.  Lanes of one instruction address distinct
// nodes and one wavefront owns the ring, so the order of additions into every
// slot is fixed by the program and the binning alone.
template <typename ValT> struct LdsAcc {
    using slot_t = double;
    __device__ static __forceinline__ void zero(slot_t& a) { a = 0.0; }
    __device__ static __forceinline__ void add(slot_t* a, ValT v) { atomicAdd(a, (double)v); }
    __device__ static __forceinline__ ValT value(const slot_t& a) { return (ValT)a; }
};

// Value of lane ex-K (row_shr by K; K = 0: this lane's own).
template <int K, typename ValT>
__device__ __forceinline__ ValT from_left(ValT v) {
    if constexpr (K == 0) return v;
    else return row_shr<K>(v);
}

// Data of the particle K lanes below: G[c] = wx_K * u_c and its z weights
// (its y weights are fetched per y offset, to keep fewer registers live).
template <int K, int KS, typename ValT, int LDA>
__device__ __forceinline__ void shift_in(const ValT (&wx)[KS], const ValT (&wz)[KS], const ValT (&u)[LDA],
                                         ValT (&G)[LDA], ValT (&WZ)[KS]) {
#pragma unroll
    for (int c = 0; c < LDA; ++c) G[c] = from_left<K>(wx[K] * u[c]);
#pragma unroll
    for (int o = 0; o < KS; ++o) WZ[o] = from_left<K>(wz[o]);
}

// Tiled deposit.  Lane = one cell of the 16 x (TY+3) window: ex = lane & 15
// along x, so each 16-lane DPP row is a run of 16 consecutive cells at fixed
// y; the window's cell rows are taken 4 at a time (one per DPP row), a
// "chunk".  In round r every lane takes the r-th particle of its cell (zero
// strength when it has none).  Lane ex finishes node ex-3 of its row, which
// receives x-offset k from the particle of lane ex-k:
//   node(ex - 3) += sum_k wx_k(ex-k) u(ex-k) wy_oy(ex-k) wz_oz(ex-k)
// The 11 inputs of each neighbour (wx_k u, wy, wz) come over by DPP row
// shifts, the 16 (y,z) sums are formed locally, and each lane adds ONE value
// per (y, z offset, component) to the ring -- 48 LDS atomics per particle for
// lda = 3 instead of 192.  The ring is stored component-major (SoA) so a row's
// 16 lanes hit consecutive 8-byte words.
//
// NW wavefronts share one ring.  With NW = 1 the wavefront takes the chunks
// of a plane in order.  With NW > 1 the tile has 2*NW chunks per plane and
// every plane runs in two phases: wavefront w takes chunk 2w in the first and
// chunk 2w+1 in the second.  Chunk q adds to node rows 4q-3 .. 4q+3, so the
// chunks of one phase touch disjoint nodes, and a barrier orders the phases:
// every node receives its additions in an order fixed by the program and the
// binning, as with one wavefront.  A taller tile (more wavefronts) re-reads
// fewer halo cells: the window is 16 x (TY+3) cells for 13 x TY nodes.
//
// TAB: the 1D weights come from the caller's tables (ppm_rmsh_remesh's
// wx1/wx2/wx3, four per particle and axis) instead of the positions.
//
// TABLE: the binning is a slot table (byte counts `bt.count`, particle r of
// cell c at bt.table[r * ncells + c]) instead of heads and a permutation: a
// cell's count and its first particle load together, one dependent load less
// before the particle's own.
//
// KS = 2 (BSP2) is the same march with a 1-cell halo: the window is 16 x
// (TY+1) cells for 15 x TY nodes, lane ex finishes node ex-1 from lanes ex
// and ex-1, cell plane cz adds to node planes cz and cz+1 (a ring of two).

template <int KS, int LDA, typename PosT, typename ValT, int TY, int NW, bool OPEN, bool TAB, bool TABLE>
__global__ void __launch_bounds__(64 * NW)
k_p3_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, BinTable bt,
          const PosT* __restrict__ xp, const ValT* __restrict__ up, FieldGeom fg, int c0, int tz,
          ValT* __restrict__ field, WTab<ValT> wt) {
    static_assert(KS == 4 || !TAB, "weight tables are M'4 only");
    constexpr int H = Stencil<KS>::H, B = Stencil<KS>::B;
    constexpr int TT = 16 - H;                     // owned nodes per tile along x
    constexpr int RP = KS;                         // ring planes
    const int lda = fg.lda;
    // per-particle data besides the strengths: the position, or (TAB) the
    // twelve 1D weights
    constexpr int NPD = TAB ? 12 : 3;
    using PD = std::conditional_t<TAB, ValT, PosT>;
    using A = LdsAcc<ValT>;
    using slot_t = typename A::slot_t;
    constexpr int NT = 64 * NW;
    constexpr int P3_PL = TT * TY;                 // nodes per ring plane
    constexpr int NQ = (TY + H) / 4;               // chunks of 4 cell rows per plane
    static_assert((TY + H) % 4 == 0, "tile height + halo must be a multiple of 4");
    static_assert(NW == 1 || NQ == 2 * NW, "with several wavefronts each takes one chunk per phase");
    constexpr int NC = NW == 1 ? NQ : 2;           // chunks per wavefront and plane
    constexpr int PLANE = P3_PL * LDA;            // slots per ring plane: [c][node]
    __shared__ slot_t ring[RP * PLANE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // nodes the tiles produce per axis: the N unique nodes of a periodic axis,
    // the N+1 real nodes 0..N of an open one
    const int Rx = m.N[0] + !m.periodic[0], Ry = m.N[1] + !m.periodic[1], Rz = m.N[2] + !m.periodic[2];
    const int ntx = (Rx + TT - 1) / TT, nty = (Ry + TY - 1) / TY;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int tx0 = (int)(tile % ntx) * TT, ty0 = (int)((tile / ntx) % nty) * TY;
    const int z0 = (int)(tile / ((int64_t)ntx * nty)) * tz;
    const int z1 = min(z0 + tz, Rz);
    const int nxo = min(TT, Rx - tx0), nyo = min(TY, Ry - ty0);
    for (int i = tid; i < RP * PLANE; i += NT) A::zero(ring[i]);

    const int ex = lane & 15, erow = lane >> 4;
    const int lx = ex - H;                          // node this lane owns after the row sum
    const bool xown = lx >= 0 && lx < nxo;
    // cell c of the window along axis d: a periodic axis wraps, an open one
    // has no cell outside 0..nbin-1 (-1).  OPEN: some axis is open (a fully
    // periodic mesh compiles without the checks).
    auto wcell = [&](int c, int d) {
        const int nb = m.nbin[d];
        if (!OPEN || m.periodic[d]) {
            c %= nb;
            return c < 0 ? c + nb : c;
        }
        return c >= 0 && c < nb ? c : -1;
    };
    // window cells start H - B below the tile's first node on x and y; cell
    // planes run from z0 - (H - B) to z1 - 1 + B
    const int cx = wcell(tx0 - (H - B) + ex, 0);
    int keyxy[NC], ey4[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int q = NW == 1 ? j : 2 * wave + j;
        int ey = q * 4 + erow;
        ey4[j] = ey;
        const int cy = wcell(ty0 - (H - B) + ey, 1);
        keyxy[j] = OPEN && (cx < 0 || cy < 0) ? -1 : cx + m.nbin[0] * cy;
    }
    __syncthreads();
    const int64_t plane_cells = (int64_t)m.nbin[0] * m.nbin[1];
    constexpr int WROW = TT * LDA;                 // values of one node row
    constexpr int WRG = NT / WROW;                 // node rows written at once
    static_assert(WRG >= 1, "a workgroup writes at least one node row at once");
    const int cz0 = z0 - (H - B), cz1 = z1 - 1 + B;

    // The cell heads of the step after next are loaded while the current
    // (plane, chunk) step computes.  (f, n) is the cell's range [f, n) of the
    // permutation, or (TABLE) its key and its count word, with t its first
    // particle (loaded even for an empty cell: the table is allocated there).
    auto heads = [&](int cz, int j, int& f, int& n, int& t) {
        f = 0;
        n = 0;
        t = 0;
        const int czw = cz <= cz1 ? wcell(cz, 2) : -1;
        if (czw >= 0 && (!OPEN || keyxy[j] >= 0)) {
            const int key = (int)(plane_cells * czw + keyxy[j]);
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
    // particles in the cell of (f, n), and its particle r < that count
    auto count_of = [&](int f, int n) {
        if constexpr (TABLE) return (int)(((unsigned)n >> count_shift(f)) & 255u);
        else return n - f;
    };
    auto entry = [&](int f, int r) {
        if constexpr (TABLE) return bt.table[(int64_t)r * bt.ncells + f];
        else return perm ? perm[f + r] : f + r;
    };
    // position (or weights) and strengths of particle p (an index into the
    // caller's arrays) when has, zeros otherwise
    auto load = [&](int p, bool has, PD (&x)[NPD], ValT (&u)[LDA]) {
#pragma unroll
        for (int d = 0; d < NPD; ++d) x[d] = PD(0);
#pragma unroll
        for (int c = 0; c < LDA; ++c) u[c] = ValT(0);
        if (has) {
            if constexpr (TAB) {
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int o = 0; o < 4; ++o) x[4 * d + o] = wt.w[d][(int64_t)p * 4 + o];
            } else {
#pragma unroll
                for (int d = 0; d < 3; ++d) x[d] = xp[(int64_t)p * 3 + d];
            }
#pragma unroll
            for (int c = 0; c < LDA; ++c) u[c] = up[(int64_t)p * lda + c0 + c];
        }
    };
    // cell offsets (INT truncation as ppm_interp_p2m.f:800-822, in the position precision)
    auto offsets = [&](const PosT (&x)[3], ValT (&fr)[3]) {
        int ci;
#pragma unroll
        for (int d = 0; d < 3; ++d) fr[d] = (ValT)cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], ci);
    };
    // the 1D weights of a particle
    auto weights = [&](const PD (&x)[NPD], ValT (&wx)[KS], ValT (&wy)[KS], ValT (&wz)[KS]) {
        if constexpr (TAB) {
#pragma unroll
            for (int o = 0; o < KS; ++o) {
                wx[o] = x[o];
                wy[o] = x[4 + o];
                wz[o] = x[8 + o];
            }
        } else {
            ValT fr[3];
            offsets(x, fr);
            tile_weights<KS, ValT>(fr[0], wx);
            tile_weights<KS, ValT>(fr[1], wy);
            tile_weights<KS, ValT>(fr[2], wz);
        }
    };
    // The permutation entry of a step's first particle is looked up one step
    // ahead (TABLE: it came with the heads), and its position and strengths at
    // the end of the step before.
    auto lookup = [&](int f, int n, int t) {
        if constexpr (TABLE) return count_of(f, n) > 0 ? t : 0;
        else return n > f ? (perm ? perm[f] : f) : 0;
    };
    int f1, n1, t1, f2, n2, t2;
    heads(cz0, 0, f1, n1, t1);
    if (NC > 1) heads(cz0, 1, f2, n2, t2);
    else heads(cz0 + 1, 0, f2, n2, t2);
    int pn = lookup(f1, n1, t1);
    // Round 0's particle of the next step is loaded at the end of a step, while
    // the wavefront waits at the phase barrier; it is not live during the rounds.
    PD xn[NPD];
    ValT un[LDA];
    load(pn, count_of(f1, n1) > 0, xn, un);

    for (int cz = cz0; cz <= cz1; ++cz) {
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const int first = f1, cnt = count_of(f1, n1);
            // take this step's heads (loaded two steps ago) before the next loads
            // are issued, so that waiting for them does not wait for those loads too
            asm volatile("" ::"v"(first), "v"(cnt) : "memory");
            f1 = f2;
            n1 = n2;
            t1 = t2;
            pn = lookup(f1, n1, t1);
            if (j + 2 < NC) heads(cz, j + 2, f2, n2, t2);
            else if (j + 2 - NC < NC) heads(cz + 1, j + 2 - NC, f2, n2, t2);
            else heads(cz + 2, 0, f2, n2, t2);
            const int ly0 = ey4[j] - H;              // node row of stencil offset oy = 0
            // one round: every lane deposits one particle of its cell (zero
            // strength where it has none)
            auto round = [&](const PD (&x)[NPD], const ValT (&u)[LDA]) {
                ValT wx[KS], wy[KS], wz[KS];
                weights(x, wx, wy, wz);
                // Lane ex finishes node ex-H, which takes x-offset k from the particle
                // of lane ex-k: fetch that particle's wx_k * u, wy and wz by DPP
                // (k = 1..H), then form the KS^2 (y,z) sums locally.
                ValT G[KS][LDA], WZ[KS][KS];
                shift_in<0, KS>(wx, wz, u, G[0], WZ[0]);
                shift_in<1, KS>(wx, wz, u, G[1], WZ[1]);
                if constexpr (KS == 4) {
                    shift_in<2, KS>(wx, wz, u, G[2], WZ[2]);
                    shift_in<3, KS>(wx, wz, u, G[3], WZ[3]);
                }
#pragma unroll
                for (int oy = 0; oy < KS; ++oy) {
                    ValT WY[KS];
                    WY[0] = wy[oy];
                    WY[1] = from_left<1>(wy[oy]);
                    if constexpr (KS == 4) {
                        WY[2] = from_left<2>(wy[oy]);
                        WY[3] = from_left<3>(wy[oy]);
                    }
                    const int ly = ly0 + oy;
                    const bool own = xown && ly >= 0 && ly < nyo;
                    const int node = lx + TT * ly;
                    // per component: fold the y weight in first (KS products),
                    // then one z contraction per z offset
#pragma unroll
                    for (int c = 0; c < LDA; ++c) {
                        ValT Y[KS];
#pragma unroll
                        for (int k = 0; k < KS; ++k) Y[k] = WY[k] * G[k][c];
#pragma unroll
                        for (int oz = 0; oz < KS; ++oz) {
                            ValT v = WZ[0][oz] * Y[0];
#pragma unroll
                            for (int k = 1; k < KS; ++k) v += WZ[k][oz] * Y[k];
                            if (own) A::add(ring + ((cz - B + oz) & (RP - 1)) * PLANE + c * P3_PL + node, v);
                        }
                    }
                }
            };
            // round 0 with the particle loaded during the previous step
            if (__ballot(cnt > 0) != 0ull) round(xn, un);
            // Later rounds load their own particles.  A round costs the same
            // however few lanes still have a particle, so once at most P3_SERIAL
            // lanes do, their remaining particles are deposited one at a time:
            // the 64 lanes take the 64 nodes of the particle's stencil (distinct
            // nodes, so the additions stay in a fixed order; with KS = 2 lanes
            // 0..7 take the 8 nodes).
            for (int r = 1;; ++r) {
                uint64_t act = __ballot(cnt > r);
                if (act == 0ull) break;
                if (__popcll(act) > P3_SERIAL) {
                    PD x[NPD];
                    ValT u[LDA];
                    const int p = cnt > r ? entry(first, r) : 0;
                    load(p, cnt > r, x, u);
                    round(x, u);
                    continue;
                }
                const int ox = lane & (KS - 1), oy = (lane / KS) & (KS - 1), oz = lane / (KS * KS);
                const bool slane = lane < KS * KS * KS;
                slot_t* pl = ring + ((cz - B + oz) & (RP - 1)) * PLANE;
                const int qrow = ey4[j] - erow;      // first cell row of the chunk
                do {
                    const int src = __builtin_ctzll(act);
                    act &= act - 1;
                    const int sf = __builtin_amdgcn_readlane(first, src);
                    const int sc = __builtin_amdgcn_readlane(cnt, src);
                    // node of this lane's stencil offset, relative to the tile
                    const int nx = (src & 15) - H + ox, ny = qrow + (src >> 4) - H + oy;
                    const bool own = slane && nx >= 0 && nx < nxo && ny >= 0 && ny < nyo;
                    for (int rr = r; rr < sc; ++rr) {
                        PD x[NPD];
                        ValT u[LDA], wyz, wxo;
                        load(entry(sf, rr), true, x, u);
                        if constexpr (TAB) {
                            wyz = x[4 + oy] * x[8 + oz];
                            wxo = x[ox];
                        } else {
                            ValT fr[3];
                            offsets(x, fr);
                            wyz = tile_weight_at<KS, ValT>(fr[1], oy) * tile_weight_at<KS, ValT>(fr[2], oz);
                            wxo = tile_weight_at<KS, ValT>(fr[0], ox);
                        }
                        if (own) {
#pragma unroll
                            for (int c = 0; c < LDA; ++c) A::add(pl + c * P3_PL + nx + TT * ny, wyz * (wxo * u[c]));
                        }
                    }
                } while (act != 0ull);
                break;
            }
            load(pn, count_of(f1, n1) > 0, xn, un);
            // end of a phase: the other phase's chunks share node rows with this one's
            if (NW > 1 || j == NC - 1) __syncthreads();
        }
        // node plane cz-B has all its cells: write it if owned, then recycle the
        // slot.  Only owned nodes ever receive additions, so an owned plane is
        // cleared by the thread that writes each node out (one thread per slot:
        // no wavefront clears a slot another has yet to read); a plane outside
        // the chunk is cleared whole.
        const int n = cz - B;
        slot_t* pl = ring + (n & (RP - 1)) * PLANE;
        if (n >= z0 && n < z1) {
            // thread (wx, wc, wy0) writes node wx, component wc of rows wy0,
            // wy0 + WRG, ... at its place in the stored box.  A stored plane
            // fits in 2^31 bytes (checked at launch): 32-bit byte offsets.  The slots are recomputed here
            // (the barrier keeps them from being hoisted and held in registers
            // through the rounds).
            int t = tid;
            asm volatile("" : "+v"(t));
            const int wy0 = t / WROW, wx = (t % WROW) / LDA, wc = t % LDA;
            const bool wact = wy0 < WRG && wx < nxo;
            const int wlds = wc * P3_PL + wx + TT * wy0;
            const int ex0 = fg.ext[0];
            const unsigned wbyte =
                (unsigned)(((ty0 + wy0 - fg.lo[1]) * ex0 + tx0 + wx - fg.lo[0]) * lda + c0 + wc) *
                (unsigned)sizeof(ValT);
            const unsigned wstep = (unsigned)(WRG * ex0 * lda) * (unsigned)sizeof(ValT);
            const char* out = (const char*)(field + (int64_t)(n - fg.lo[2]) * ex0 * fg.ext[1] * lda);
            if (wact) {
#pragma unroll
                for (int k = 0; k < (TY + WRG - 1) / WRG; ++k) {
                    if (wy0 + k * WRG < nyo) {
                        slot_t& a = pl[wlds + k * WRG * TT];
                        const ValT v = (ValT)A::value(a);
                        A::zero(a);
                        *(ValT*)(out + (wbyte + (unsigned)k * wstep)) = v;
                    }
                }
            }
        } else {
            for (int i = tid; i < PLANE; i += NT) A::zero(pl[i]);
        }
        __syncthreads();
    }
}

static int p3_rows();

// Per axis the stored box is either the compact layout of a periodic axis
// (the N unique nodes, lo = 0, ext = N) or holds the whole real range 0..N
// (lo <= 0 <= N <= lo+ext-1: the reference's 1-g..Nm+g box is lo = -g,
// ext = N+1+2g), on periodic and open axes alike.
bool p2m3d_fast_applicable(const DevMesh& m, int kernel, int prec, const FieldGeom& fg) {
    if (m.dim != 3 || (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)) return false;
    for (int d = 0; d < 3; ++d) {
        if (m.N[d] < 4) return false;
        const bool compact = m.periodic[d] && fg.lo[d] == 0 && fg.ext[d] == m.N[d];
        const bool boxed = fg.lo[d] <= 0 && fg.lo[d] + fg.ext[d] - 1 >= m.N[d];
        if (!compact && !boxed) return false;
    }
    // the write-out addresses a stored plane with 32-bit byte offsets
    const int64_t vsize = prec == PPMG_PREC_F64 ? 8 : 4;
    return (int64_t)fg.ext[0] * fg.ext[1] * fg.lda * vsize < ((int64_t)1 << 31);
}

// The deposit (and the stored box's ghost shell) reads a slot-table binning
// on a periodic mesh, in either layout, with the default tile.
bool p2m3d_table_applicable(const DevMesh& m, int kernel, int prec, const FieldGeom& fg) {
    if (!p2m3d_fast_applicable(m, kernel, prec, fg)) return false;
    if (!(m.periodic[0] && m.periodic[1] && m.periodic[2])) return false;
    return kernel == PPMG_KERNEL_BSP2 || p3_rows() == P3_TY;
}

// One launch of the deposit for components c0 .. c0+LL-1 with tiles of
// (16 - H) x TY nodes and NW wavefronts per tile.  The z chunk is chosen from
// the device's resident workgroup count (queried once per instantiation):
// each chunk marches H extra cell planes.
template <int KS, int LL, typename PosT, typename ValT, int TY, int NW, bool OPEN, bool TAB, bool TABLE = false>
static void launch_p3_shape(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                            const FieldGeom& fg, int c0, ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    constexpr int H = Stencil<KS>::H, TT = 16 - H;
    static std::atomic<int> slots{0};
    if (slots.load() == 0)
        slots.store(resident_workgroups(k_p3_tile<KS, LL, PosT, ValT, TY, NW, OPEN, TAB, TABLE>, 64 * NW));
    const int Rx = m.N[0] + !m.periodic[0], Ry = m.N[1] + !m.periodic[1], Rz = m.N[2] + !m.periodic[2];
    const int tx = (Rx + TT - 1) / TT, ty = (Ry + TY - 1) / TY;
    // crowded cells (more than 1/1024 of the cells hold nine or more
    // particles): shorter chunks keep the workgroups of crowded tiles short
    // (clustered 512^3: 13.8 against 15.6 ms); the lattice keeps the chunk
    // that costs the least halo (4.56 against 4.82 ms with 16 planes)
    const int64_t ncells = (int64_t)m.nbin[0] * m.nbin[1] * m.nbin[2];
    const int tz = plan->crowded_cells > ncells / 1024 ? pick_z_chunk_balanced((int64_t)tx * ty, Rz, slots.load(), 16)
                                                       : pick_z_chunk((int64_t)tx * ty, Rz, slots.load(), H);
    dim3 grid((unsigned)((int64_t)tx * ty * ((Rz + tz - 1) / tz)));
    const BinTable bt{plan->count.as<uint32_t>(), plan->table.as<int>(), plan->ncells};
    hipLaunchKernelGGL((k_p3_tile<KS, LL, PosT, ValT, TY, NW, OPEN, TAB, TABLE>), grid, dim3(64 * NW), 0, s, m,
                       plan->start.as<int>(), perm, bt, xp, up, fg, c0, tz, field, wt);
}

// Tile height: PPMG_P2M_ROWS (9, 13, 21 or 29 node rows: 1, 2, 3 or 4
// wavefronts per tile) overrides the default of 21, for tuning.  Measured on
// the 512^3 advected lattice (lda 3, FP64): 5.90, 5.49, 5.39 and 5.59 ms --
// taller tiles re-read fewer halo cells until the phase barriers and the
// larger ring (fewer resident workgroups) cost more than they save.
static int p3_rows() {
    const char* e = std::getenv("PPMG_P2M_ROWS");
    const int r = e ? std::atoi(e) : 0;
    return r == 9 || r == 13 || r == 21 || r == 29 ? r : P3_TY;
}

template <int KS, int LL, typename PosT, typename ValT, bool OPEN, bool TAB>
static void launch_p3_open(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                           const FieldGeom& fg, int c0, ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    if constexpr (KS == 2) {
        // BSP2: 15 x 23 node tiles (24 cell rows, three wavefronts)
        launch_p3_shape<KS, LL, PosT, ValT, 23, 3, OPEN, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
    } else if constexpr (TAB) {
        // the remesh path takes the default tile only
        launch_p3_shape<KS, LL, PosT, ValT, 21, 3, OPEN, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
    } else {
#define SHAPE(TY, NW) \
    launch_p3_shape<KS, LL, PosT, ValT, TY, NW, OPEN, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s)
        switch (p3_rows()) {
            case 9: SHAPE(9, 1); break;
            case 13: SHAPE(13, 2); break;
            case 29: SHAPE(29, 4); break;
            default: SHAPE(21, 3); break;
        }
#undef SHAPE
    }
}

template <int KS, int LL, typename PosT, typename ValT, bool TAB>
static void launch_p3(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                      const FieldGeom& fg, int c0, ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    if constexpr (!TAB) {
        // a slot table (p2m3d_table_applicable: periodic, default tile)
        if (plan->bin_table) {
            launch_p3_shape<KS, LL, PosT, ValT, KS == 2 ? 23 : P3_TY, 3, false, false, true>(plan, m, xp, up, perm, fg,
                                                                                            c0, field, wt, s);
            return;
        }
    }
    if (m.periodic[0] && m.periodic[1] && m.periodic[2])
        launch_p3_open<KS, LL, PosT, ValT, false, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
    else
        launch_p3_open<KS, LL, PosT, ValT, true, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
}

template <int KS, typename PosT, typename ValT>
static int p2m3d_typed(ppmg_plan* plan, const DevMesh& m, int64_t np, const PosT* xp, const FieldGeom& fg,
                       const ValT* up, ValT* field, const ValT* const* wtab, hipStream_t s) {
    mark(plan, s);
    const int lda = fg.lda;
    WTab<ValT> wt{{nullptr, nullptr, nullptr}};
    if (wtab)
        for (int d = 0; d < 3; ++d) wt.w[d] = wtab[d];
    (void)np;
    // the slot table is read on a periodic mesh with the default tile, without
    // weight tables; elsewhere the heads and permutation are built from it
    if (plan->bin_table && (wtab || !(m.periodic[0] && m.periodic[1] && m.periodic[2]) || (KS == 4 && p3_rows() != P3_TY)))
        if (materialize_heads(plan, s)) return 1;
    const int* pk = tile_perm(plan);
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL)                                                                            \
    do {                                                                                  \
        if constexpr (KS == 4) {                                                          \
            if (wtab) launch_p3<KS, LL, PosT, ValT, true>(plan, m, xp, up, pk, fg, c0, field, wt, s);   \
            else launch_p3<KS, LL, PosT, ValT, false>(plan, m, xp, up, pk, fg, c0, field, wt, s);       \
        } else {                                                                          \
            launch_p3<KS, LL, PosT, ValT, false>(plan, m, xp, up, pk, fg, c0, field, wt, s);            \
        }                                                                                 \
    } while (0)
        if (L == 1) TL(1);
        else if (L == 2) TL(2);
        else if (L == 3) TL(3);
        else TL(4);
#undef TL
        PPMG_HIP(hipGetLastError());
    }
    if (fill_stored_box<3, KS, PosT, ValT>(plan, m, xp, up, fg, field, wt, s)) return 1;
    PPMG_HIP(hipGetLastError());
    mark(plan, s);
    return 0;
}

int p2m3d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp,
               const FieldGeom& fg, const void* up, void* field, const void* const* wtab, hipStream_t s) {
#define TYPED(KS)                                                                                            \
    do {                                                                                                     \
        if (prec == PPMG_PREC_F64)                                                                           \
            return p2m3d_typed<KS, double, double>(plan, m, np, (const double*)xp, fg, (const double*)up,   \
                                                   (double*)field, (const double* const*)wtab, s);          \
        if (prec == PPMG_PREC_F32C)                                                                          \
            return p2m3d_typed<KS, double, float>(plan, m, np, (const double*)xp, fg, (const float*)up,     \
                                                  (float*)field, (const float* const*)wtab, s);             \
        return p2m3d_typed<KS, float, float>(plan, m, np, (const float*)xp, fg, (const float*)up,           \
                                             (float*)field, (const float* const*)wtab, s);                  \
    } while (0)
    if (kernel == PPMG_KERNEL_BSP2) {
        if (wtab) return PPMG_ERR_ARGUMENT;   // weight tables are M'4 only
        TYPED(2);
    }
    TYPED(4);
#undef TYPED
}

}  // namespace ppmg
