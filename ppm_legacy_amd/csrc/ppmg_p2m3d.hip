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
template <int K, typename ValT, int LDA>
__device__ __forceinline__ void shift_in(const ValT (&wx)[4], const ValT (&wz)[4], const ValT (&u)[LDA],
                                         ValT (&G)[LDA], ValT (&WZ)[4]) {
#pragma unroll
    for (int c = 0; c < LDA; ++c) G[c] = from_left<K>(wx[K] * u[c]);
#pragma unroll
    for (int o = 0; o < 4; ++o) WZ[o] = from_left<K>(wz[o]);
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
template <typename ValT> struct WTab {
    const ValT* w[3];
};

template <int LDA, typename PosT, typename ValT, int TY, int NW, bool OPEN, bool TAB>
__global__ void __launch_bounds__(64 * NW)
k_p3_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, const PosT* __restrict__ xp,
          const ValT* __restrict__ up, FieldGeom fg, int c0, int tz, ValT* __restrict__ field, WTab<ValT> wt) {
    const int lda = fg.lda;
    // per-particle data besides the strengths: the position, or (TAB) the
    // twelve 1D weights
    constexpr int NPD = TAB ? 12 : 3;
    using PD = std::conditional_t<TAB, ValT, PosT>;
    using A = LdsAcc<ValT>;
    using slot_t = typename A::slot_t;
    constexpr int NT = 64 * NW;
    constexpr int P3_PL = P3_TT * TY;              // nodes per ring plane
    constexpr int NQ = (TY + 3) / 4;               // chunks of 4 cell rows per plane
    static_assert((TY + 3) % 4 == 0, "tile height + 3 must be a multiple of 4");
    static_assert(NW == 1 || NQ == 2 * NW, "with several wavefronts each takes one chunk per phase");
    constexpr int NC = NW == 1 ? NQ : 2;           // chunks per wavefront and plane
    constexpr int PLANE = P3_PL * LDA;            // slots per ring plane: [c][node]
    __shared__ slot_t ring[4 * PLANE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // nodes the tiles produce per axis: the N unique nodes of a periodic axis,
    // the N+1 real nodes 0..N of an open one
    const int Rx = m.N[0] + !m.periodic[0], Ry = m.N[1] + !m.periodic[1], Rz = m.N[2] + !m.periodic[2];
    const int ntx = (Rx + P3_TT - 1) / P3_TT, nty = (Ry + TY - 1) / TY;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int tx0 = (int)(tile % ntx) * P3_TT, ty0 = (int)((tile / ntx) % nty) * TY;
    const int z0 = (int)(tile / ((int64_t)ntx * nty)) * tz;
    const int z1 = min(z0 + tz, Rz);
    const int nxo = min(P3_TT, Rx - tx0), nyo = min(TY, Ry - ty0);
    for (int i = tid; i < 4 * PLANE; i += NT) A::zero(ring[i]);

    const int ex = lane & 15, erow = lane >> 4;
    const int lx = ex - 3;                          // node this lane owns after the row sum
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
    const int cx = wcell(tx0 - 2 + ex, 0);
    int keyxy[NC], ey4[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int q = NW == 1 ? j : 2 * wave + j;
        int ey = q * 4 + erow;
        ey4[j] = ey;
        const int cy = wcell(ty0 - 2 + ey, 1);
        keyxy[j] = OPEN && (cx < 0 || cy < 0) ? -1 : cx + m.nbin[0] * cy;
    }
    __syncthreads();
    const int64_t plane_cells = (int64_t)m.nbin[0] * m.nbin[1];
    constexpr int WROW = P3_TT * LDA;              // values of one node row
    constexpr int WRG = NT / WROW;                 // node rows written at once
    static_assert(WRG >= 1, "a workgroup writes at least one node row at once");

    // The cell heads of the step after next are loaded while the current
    // (plane, chunk) step computes.
    auto heads = [&](int cz, int j, int& f, int& n) {
        f = 0;
        n = 0;
        const int czw = cz <= z1 ? wcell(cz, 2) : -1;
        if (czw >= 0 && (!OPEN || keyxy[j] >= 0)) {
            const int key = (int)(plane_cells * czw + keyxy[j]);
            f = start[key];
            n = start[key + 1];   // end, subtracted where used: the load stays in flight
        }
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
    auto weights = [&](const PD (&x)[NPD], ValT (&wx)[4], ValT (&wy)[4], ValT (&wz)[4]) {
        if constexpr (TAB) {
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                wx[o] = x[o];
                wy[o] = x[4 + o];
                wz[o] = x[8 + o];
            }
        } else {
            ValT fr[3];
            offsets(x, fr);
            mp4_factored<ValT>(fr[0], wx);
            mp4_factored<ValT>(fr[1], wy);
            mp4_factored<ValT>(fr[2], wz);
        }
    };
    // The permutation entry of a step's first particle is looked up one step
    // ahead, and its position and strengths at the end of the step before.
    auto lookup = [&](int f, int n) { return n > f ? (perm ? perm[f] : f) : 0; };
    int f1, n1, f2, n2;
    heads(z0 - 2, 0, f1, n1);
    if (NC > 1) heads(z0 - 2, 1, f2, n2);
    else heads(z0 - 1, 0, f2, n2);
    int pn = lookup(f1, n1);
    // Round 0's particle of the next step is loaded at the end of a step, while
    // the wavefront waits at the phase barrier; it is not live during the rounds.
    PD xn[NPD];
    ValT un[LDA];
    load(pn, n1 > f1, xn, un);

    for (int cz = z0 - 2; cz <= z1; ++cz) {
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const int first = f1, cnt = n1 - f1;   // n: end of the cell's range
            // take this step's heads (loaded two steps ago) before the next loads
            // are issued, so that waiting for them does not wait for those loads too
            asm volatile("" ::"v"(first), "v"(cnt) : "memory");
            f1 = f2;
            n1 = n2;
            pn = lookup(f1, n1);
            if (j + 2 < NC) heads(cz, j + 2, f2, n2);
            else if (j + 2 - NC < NC) heads(cz + 1, j + 2 - NC, f2, n2);
            else heads(cz + 2, 0, f2, n2);
            const int ly0 = ey4[j] - 3;              // node row of stencil offset oy = 0
            // one round: every lane deposits one particle of its cell (zero
            // strength where it has none)
            auto round = [&](const PD (&x)[NPD], const ValT (&u)[LDA]) {
                ValT wx[4], wy[4], wz[4];
                weights(x, wx, wy, wz);
                // Lane ex finishes node ex-3, which takes x-offset k from the particle
                // of lane ex-k: fetch that particle's wx_k * u, wy and wz by DPP
                // (k = 1..3), then form the 16 (y,z) sums locally.
                ValT G[4][LDA], WZ[4][4];
                shift_in<0>(wx, wz, u, G[0], WZ[0]);
                shift_in<1>(wx, wz, u, G[1], WZ[1]);
                shift_in<2>(wx, wz, u, G[2], WZ[2]);
                shift_in<3>(wx, wz, u, G[3], WZ[3]);
#pragma unroll
                for (int oy = 0; oy < 4; ++oy) {
                    const ValT WY[4] = {from_left<0>(wy[oy]), from_left<1>(wy[oy]), from_left<2>(wy[oy]),
                                        from_left<3>(wy[oy])};
                    const int ly = ly0 + oy;
                    const bool own = xown && ly >= 0 && ly < nyo;
                    const int node = lx + P3_TT * ly;
                    // per component: fold the y weight in first (4 products),
                    // then one z contraction per z offset
#pragma unroll
                    for (int c = 0; c < LDA; ++c) {
                        const ValT H[4] = {WY[0] * G[0][c], WY[1] * G[1][c], WY[2] * G[2][c], WY[3] * G[3][c]};
#pragma unroll
                        for (int oz = 0; oz < 4; ++oz) {
                            ValT v = WZ[0][oz] * H[0];
                            v += WZ[1][oz] * H[1];
                            v += WZ[2][oz] * H[2];
                            v += WZ[3][oz] * H[3];
                            if (own) A::add(ring + ((cz - 1 + oz) & 3) * PLANE + c * P3_PL + node, v);
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
            // nodes, so the additions stay in a fixed order).
            for (int r = 1;; ++r) {
                uint64_t act = __ballot(cnt > r);
                if (act == 0ull) break;
                if (__popcll(act) > P3_SERIAL) {
                    PD x[NPD];
                    ValT u[LDA];
                    const int p = cnt > r ? (perm ? perm[first + r] : first + r) : 0;
                    load(p, cnt > r, x, u);
                    round(x, u);
                    continue;
                }
                const int ox = lane & 3, oy = (lane >> 2) & 3, oz = lane >> 4;
                slot_t* pl = ring + ((cz - 1 + oz) & 3) * PLANE;
                const int qrow = ey4[j] - erow;      // first cell row of the chunk
                do {
                    const int src = __builtin_ctzll(act);
                    act &= act - 1;
                    const int sf = __builtin_amdgcn_readlane(first, src);
                    const int sc = __builtin_amdgcn_readlane(cnt, src);
                    // node of this lane's stencil offset, relative to the tile
                    const int nx = (src & 15) - 3 + ox, ny = qrow + (src >> 4) - 3 + oy;
                    const bool own = nx >= 0 && nx < nxo && ny >= 0 && ny < nyo;
                    for (int rr = r; rr < sc; ++rr) {
                        PD x[NPD];
                        ValT u[LDA], wyz, wxo;
                        load(perm ? perm[sf + rr] : sf + rr, true, x, u);
                        if constexpr (TAB) {
                            wyz = x[4 + oy] * x[8 + oz];
                            wxo = x[ox];
                        } else {
                            ValT fr[3];
                            offsets(x, fr);
                            wyz = mp4_at<ValT>(fr[1], oy) * mp4_at<ValT>(fr[2], oz);
                            wxo = mp4_at<ValT>(fr[0], ox);
                        }
                        if (own) {
#pragma unroll
                            for (int c = 0; c < LDA; ++c) A::add(pl + c * P3_PL + nx + P3_TT * ny, wyz * (wxo * u[c]));
                        }
                    }
                } while (act != 0ull);
                break;
            }
            load(pn, n1 > f1, xn, un);
            // end of a phase: the other phase's chunks share node rows with this one's
            if (NW > 1 || j == NC - 1) __syncthreads();
        }
        // node plane cz-1 has all its cells: write it if owned, then recycle the
        // slot.  Only owned nodes ever receive additions, so an owned plane is
        // cleared by the thread that writes each node out (one thread per slot:
        // no wavefront clears a slot another has yet to read); a plane outside
        // the chunk is cleared whole.
        const int n = cz - 1;
        slot_t* pl = ring + (n & 3) * PLANE;
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
            const int wlds = wc * P3_PL + wx + P3_TT * wy0;
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
                        slot_t& a = pl[wlds + k * WRG * P3_TT];
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

// Per axis the stored box is either the compact layout of a periodic axis
// (the N unique nodes, lo = 0, ext = N) or holds the whole real range 0..N
// (lo <= 0 <= N <= lo+ext-1: the reference's 1-g..Nm+g box is lo = -g,
// ext = N+1+2g), on periodic and open axes alike.
bool p2m3d_fast_applicable(const DevMesh& m, int kernel, int prec, const FieldGeom& fg) {
    if (m.dim != 3 || kernel != PPMG_KERNEL_MP4) return false;
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

// One launch of the deposit for components c0 .. c0+LL-1 with tiles of
// 13 x TY nodes and NW wavefronts per tile.  The z chunk is chosen from the
// device's resident workgroup count (queried once per instantiation): each
// chunk marches 3 extra cell planes.
template <int LL, typename PosT, typename ValT, int TY, int NW, bool OPEN, bool TAB>
static void launch_p3_shape(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                            const FieldGeom& fg, int c0, ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    static std::atomic<int> slots{0};
    if (slots.load() == 0) slots.store(resident_workgroups(k_p3_tile<LL, PosT, ValT, TY, NW, OPEN, TAB>, 64 * NW));
    const int Rx = m.N[0] + !m.periodic[0], Ry = m.N[1] + !m.periodic[1], Rz = m.N[2] + !m.periodic[2];
    const int tx = (Rx + P3_TT - 1) / P3_TT, ty = (Ry + TY - 1) / TY;
    const int tz = pick_z_chunk((int64_t)tx * ty, Rz, slots.load(), 3);
    dim3 grid((unsigned)((int64_t)tx * ty * ((Rz + tz - 1) / tz)));
    hipLaunchKernelGGL((k_p3_tile<LL, PosT, ValT, TY, NW, OPEN, TAB>), grid, dim3(64 * NW), 0, s, m,
                       plan->start.as<int>(), perm, xp, up, fg, c0, tz, field, wt);
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

template <int LL, typename PosT, typename ValT, bool OPEN, bool TAB>
static void launch_p3_open(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                           const FieldGeom& fg, int c0, ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    if constexpr (TAB) {
        // the remesh path takes the default tile only
        launch_p3_shape<LL, PosT, ValT, 21, 3, OPEN, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
    } else {
#define SHAPE(TY, NW) launch_p3_shape<LL, PosT, ValT, TY, NW, OPEN, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s)
        switch (p3_rows()) {
            case 9: SHAPE(9, 1); break;
            case 13: SHAPE(13, 2); break;
            case 29: SHAPE(29, 4); break;
            default: SHAPE(21, 3); break;
        }
#undef SHAPE
    }
}

template <int LL, typename PosT, typename ValT, bool TAB>
static void launch_p3(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const int* perm,
                      const FieldGeom& fg, int c0, ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    if (m.periodic[0] && m.periodic[1] && m.periodic[2])
        launch_p3_open<LL, PosT, ValT, false, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
    else
        launch_p3_open<LL, PosT, ValT, true, TAB>(plan, m, xp, up, perm, fg, c0, field, wt, s);
}

// Stored nodes besides the unique ones, for axes whose stored box holds the
// whole real range 0..N.  Per axis: R = stored real nodes (N compact, N+1
// boxed), O = ext - R stored nodes outside 0..N.

// Real nodes with index N on a boxed periodic axis (n.d = N; -1 on other
// axes): the fold leaves node N equal to node 0, so they take the value the
// tiles wrote there.  Enumerated as the face i = N, then j = N (i < N), then
// k = N (i, j < N), over the real nodes (r per axis, b = 1 where node N is a
// copy).
template <typename ValT>
__global__ void __launch_bounds__(256) k_p3_dup(FieldGeom fg, int3 n, int3 r, int3 b, ValT* __restrict__ field,
                                                int64_t ndup) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ndup) return;
    const int64_t fx = b.x ? (int64_t)r.y * r.z : 0;
    const int64_t fy = b.y ? (int64_t)(r.x - b.x) * r.z : 0;
    int i, j, k;
    if (t < fx) {
        i = n.x;
        j = (int)(t % r.y);
        k = (int)(t / r.y);
    } else if (t < fx + fy) {
        const int64_t q = t - fx;
        i = (int)(q % (r.x - b.x));
        j = n.y;
        k = (int)(q / (r.x - b.x));
    } else {
        const int64_t q = t - fx - fy;
        i = (int)(q % (r.x - b.x));
        j = (int)((q / (r.x - b.x)) % (r.y - b.y));
        k = n.z;
    }
    auto at = [&](int x, int y, int z) {
        return (((int64_t)(z - fg.lo[2]) * fg.ext[1] + (y - fg.lo[1])) * fg.ext[0] + (x - fg.lo[0])) * fg.lda;
    };
    const int64_t dst = at(i, j, k);
    const int64_t src = at(i == n.x ? 0 : i, j == n.y ? 0 : j, k == n.z ? 0 : k);
    for (int c = 0; c < fg.lda; ++c) field[dst + c] = field[src + c];
}

// Stored nodes outside 0..N on some axis (ghost layers): the raw sum of the
// particles whose stencil reaches the node with no periodic wrap on any axis,
// which is what ppm_map_field_ghost_put leaves in ghost nodes (the general
// path's rule for nodes that are not real).  Enumerated as the planes outside
// 0..N in z, then the rows outside in y (z real), then the nodes outside in x
// (y, z real).  Weights and products in the value precision, sums in FP64;
// with weight tables (wt.w[0] set) the weights are the tables'.
template <typename PosT, typename ValT, int LL>
__global__ void __launch_bounds__(256) k_p3_shell(DevMesh m, FieldGeom fg, int3 r, const int* __restrict__ start,
                                                  const int* __restrict__ perm, const PosT* __restrict__ xp,
                                                  const ValT* __restrict__ up, int c0, ValT* __restrict__ field,
                                                  int64_t nshell, WTab<ValT> wt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nshell) return;
    const int e0 = fg.ext[0], e1 = fg.ext[1], e2 = fg.ext[2];
    const int o0 = e0 - r.x, o1 = e1 - r.y, o2 = e2 - r.z;
    // q-th stored node of axis d outside the real range: below 0, then above N
    auto outer = [&](int d, int q) { return q < -fg.lo[d] ? fg.lo[d] + q : m.N[d] + 1 + (q + fg.lo[d]); };
    const int64_t ra = (int64_t)o2 * e1 * e0, rb = (int64_t)r.z * o1 * e0;
    int n[3];
    if (t < ra) {
        n[0] = fg.lo[0] + (int)(t % e0);
        n[1] = fg.lo[1] + (int)((t / e0) % e1);
        n[2] = outer(2, (int)(t / ((int64_t)e0 * e1)));
    } else if (t < ra + rb) {
        const int64_t q = t - ra;
        n[0] = fg.lo[0] + (int)(q % e0);
        n[1] = outer(1, (int)((q / e0) % o1));
        n[2] = (int)(q / ((int64_t)e0 * o1));
    } else {
        const int64_t q = t - ra - rb;
        n[0] = outer(0, (int)(q % o0));
        n[1] = (int)((q / o0) % r.y);
        n[2] = (int)(q / ((int64_t)o0 * r.y));
    }
    double acc[LL];
#pragma unroll
    for (int c = 0; c < LL; ++c) acc[c] = 0.0;
    for (int oz = 0; oz < 4; ++oz) {
        const int cz = n[2] + 1 - oz;
        if (cz < 0 || cz >= m.nbin[2]) continue;
        for (int oy = 0; oy < 4; ++oy) {
            const int cy = n[1] + 1 - oy;
            if (cy < 0 || cy >= m.nbin[1]) continue;
            for (int ox = 0; ox < 4; ++ox) {
                const int cx = n[0] + 1 - ox;
                if (cx < 0 || cx >= m.nbin[0]) continue;
                const int key = cx + m.nbin[0] * (cy + m.nbin[1] * cz);
                const int b = start[key], e = start[key + 1];
                for (int q = b; q < e; ++q) {
                    const int p = perm ? perm[q] : q;
                    ValT w;
                    if (wt.w[0]) {
                        w = (wt.w[0][(int64_t)p * 4 + ox] * wt.w[1][(int64_t)p * 4 + oy]) * wt.w[2][(int64_t)p * 4 + oz];
                    } else {
                        ValT fr[3];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            int ci;
                            fr[d] = (ValT)cell_offset<PosT>(xp[(int64_t)p * 3 + d], (PosT)m.minv[d], (PosT)m.dxi[d], ci);
                        }
                        w = (mp4_at<ValT>(fr[0], ox) * mp4_at<ValT>(fr[1], oy)) * mp4_at<ValT>(fr[2], oz);
                    }
#pragma unroll
                    for (int c = 0; c < LL; ++c) acc[c] += (double)(w * up[(int64_t)p * fg.lda + c0 + c]);
                }
            }
        }
    }
    ValT* out = field + (((int64_t)(n[2] - fg.lo[2]) * e1 + (n[1] - fg.lo[1])) * e0 + (n[0] - fg.lo[0])) * fg.lda + c0;
#pragma unroll
    for (int c = 0; c < LL; ++c) out[c] = (ValT)acc[c];
}

template <typename PosT, typename ValT>
static int p2m3d_typed(ppmg_plan* plan, const DevMesh& m, int64_t np, const PosT* xp, const FieldGeom& fg,
                       const ValT* up, ValT* field, const ValT* const* wtab, hipStream_t s) {
    mark(plan, s);
    const int lda = fg.lda;
    WTab<ValT> wt{{nullptr, nullptr, nullptr}};
    if (wtab)
        for (int d = 0; d < 3; ++d) wt.w[d] = wtab[d];
    (void)np;
    const int* pk = tile_perm(plan);
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL)                                                                            \
    do {                                                                                  \
        if (wtab) launch_p3<LL, PosT, ValT, true>(plan, m, xp, up, pk, fg, c0, field, wt, s);   \
        else launch_p3<LL, PosT, ValT, false>(plan, m, xp, up, pk, fg, c0, field, wt, s);       \
    } while (0)
        if (L == 1) TL(1);
        else if (L == 2) TL(2);
        else if (L == 3) TL(3);
        else TL(4);
#undef TL
        PPMG_HIP(hipGetLastError());
    }
    // the rest of a stored box that holds the real range: node N of periodic
    // axes (open axes' node N is real and came from the tiles), then ghosts
    int R[3], B[3], D[3];
    for (int d = 0; d < 3; ++d) {
        const bool compact = m.periodic[d] && fg.lo[d] == 0 && fg.ext[d] == m.N[d];
        B[d] = m.periodic[d] && !compact;
        R[d] = compact ? m.N[d] : m.N[d] + 1;
        D[d] = B[d] ? m.N[d] : -1;
    }
    const int3 n3 = make_int3(D[0], D[1], D[2]), r3 = make_int3(R[0], R[1], R[2]),
               b3 = make_int3(B[0], B[1], B[2]);
    const int64_t ndup = (int64_t)R[0] * R[1] * R[2] - (int64_t)(R[0] - B[0]) * (R[1] - B[1]) * (R[2] - B[2]);
    if (ndup > 0)
        hipLaunchKernelGGL(k_p3_dup<ValT>, dim3((unsigned)((ndup + 255) / 256)), dim3(256), 0, s, fg, n3, r3, b3,
                           field, ndup);
    const int64_t nshell = (int64_t)fg.ext[0] * fg.ext[1] * fg.ext[2] - (int64_t)R[0] * R[1] * R[2];
    if (nshell > 0) {
        const unsigned g = (unsigned)((nshell + 255) / 256);
        for (int c0 = 0; c0 < lda; c0 += 4) {
            const int L = lda - c0 < 4 ? lda - c0 : 4;
#define SH(LL) hipLaunchKernelGGL((k_p3_shell<PosT, ValT, LL>), dim3(g), dim3(256), 0, s, m, fg, r3, \
                                  plan->start.as<int>(), tile_perm(plan), xp, up, c0, field, nshell, wt)
            if (L == 1) SH(1);
            else if (L == 2) SH(2);
            else if (L == 3) SH(3);
            else SH(4);
#undef SH
        }
    }
    PPMG_HIP(hipGetLastError());
    mark(plan, s);
    return 0;
}

int p2m3d_fast(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp, const FieldGeom& fg,
               const void* up, void* field, const void* const* wtab, hipStream_t s) {
    if (prec == PPMG_PREC_F64)
        return p2m3d_typed<double, double>(plan, m, np, (const double*)xp, fg, (const double*)up, (double*)field,
                                           (const double* const*)wtab, s);
    if (prec == PPMG_PREC_F32C)
        return p2m3d_typed<double, float>(plan, m, np, (const double*)xp, fg, (const float*)up, (float*)field,
                                          (const float* const*)wtab, s);
    return p2m3d_typed<float, float>(plan, m, np, (const float*)xp, fg, (const float*)up, (float*)field,
                                     (const float* const*)wtab, s);
}

}  // namespace ppmg
