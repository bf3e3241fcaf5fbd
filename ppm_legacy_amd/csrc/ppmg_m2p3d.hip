// ppmg_m2p3d.hip -- 3D M'4 M2P fast path for gfx950: tiled z-march gather
// with DPP row exchange.
//
// Replaces the gather loop of ppm_interp_m2p (ppm_interp_m2p.f:895-1240) for
// 3D M'4 on periodic and open axes, reading the compact layout or a stored box;
// the periodic ghost fill the reference gets from ppm_map_field_ghost
// (ghost_get) is the node map of the planes loaded around the tile
// (stored_node).
//
// A workgroup (4 wavefronts, 16 rows of 16 lanes) owns the particles of a
// 13 x 16 cell tile in x,y over 32 cell planes in z.  The 16 x 19 nodes those
// cells read are kept for 4 node planes in an LDS ring (component-major, rows
// of 16 nodes): cell plane cz reads node planes cz-1 .. cz+2, after which
// plane cz-1 is replaced by plane cz+3, loaded with coalesced reads.
//
// Lane mapping: lane ex of row `row` stands for node column tx0-1+ex and for
// cell (tx0+ex-1, ty0+row); lanes 1..13 carry the particles of their cell, one
// particle per round (particles are binned, so a cell's particles are
// contiguous in the permutation).  Instead of every particle reading its
// 4x4x4 stencil (192 scattered LDS reads for lda = 3), every lane reads ONE
// 4x4 (y,z) column of its node column -- 16 lanes read 16 consecutive words,
// free of bank conflicts -- and contracts it with the (y,z) weights of the
// four particles whose x-stencil covers that column; those weights come from
// lanes ex+1, ex, ex-1, ex-2 by DPP row shifts, and the four partial sums go
// back to their particles by the opposite shifts, where the x weights finish
// the sum:
//   u(P) = wx0(P) C0(P-1) + wx1(P) C1(P) + wx2(P) C2(P+1) + wx3(P) C3(P+2),
//   Ck(L) = sum_{oy,oz} wy_oy(L+1-k) wz_oz(L+1-k) f(L, oy, oz).
// Output order is the caller's particle order (written through the
// permutation).
#include <atomic>

#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

constexpr int M3_TY = 16;               // owned cell rows per tile (one 16-lane row each)
constexpr int M3_THREADS = 16 * M3_TY;  // 256

// Tile geometry for a stencil of KS nodes per axis (4: M'4, 2: BSP2): a
// particle of cell c reads nodes c - B .. c - B + KS - 1.  Lane ex of a row
// stands for node column tx0 - B + ex; the cells of lanes B .. B + TX - 1
// carry particles.  The ring holds KS node planes of RY rows of 16 nodes.
template <int KS> struct M3Geom {
    static constexpr int B = KS == 4 ? 1 : 0;
    static constexpr int TX = 17 - KS;            // owned cells per tile along x (13 or 15)
    static constexpr int RY = M3_TY + KS - 1;     // node rows in the ring
    static constexpr int PL = RY * 16;            // node slots per ring plane and component
};

// Stored node read as stored: v itself on a stored box (clamped into it for
// the ring columns no stencil reads); a compact periodic axis wraps.
template <bool OPEN>
__device__ __forceinline__ int raw_node(const DevMesh& m, const FieldGeom& fg, int d, int v) {
    if (fg.swap[d] || !m.periodic[d]) return min(max(v, fg.lo[d]), fg.lo[d] + fg.ext[d] - 1);
    return stored_node<OPEN>(m, fg, d, v);
}

// WPE: minimum wavefronts per SIMD the register allocation must allow: 4 (at
// most 128 VGPRs) where that needs no spills, else 3 (M'4 with four
// components or the MIX form).  At 512^3, lda 3, FP64 the gather takes 2.94 ms
// with 4 wavefronts against 3.18 ms with 3 (136 VGPRs unconstrained).
// MIX: some open axis beside a periodic axis stored as a box, where whether a
// node takes the periodic node map depends on all three of its coordinates.
// TABLE: the binning is a slot table (byte counts, particle r of cell c at
// bt.table[r * ncells + c]) instead of heads and a permutation.
// LAT: the binning is taken to be a lattice in cell order (particle c in cell
// c, one per cell) without having been computed: every particle's cell key
// is checked against its cell, and a mismatch sets *lat_err (the caller then
// bins and gathers again).
template <int KS, int LDA, typename PosT, typename ValT, typename AccT, int WPE, bool OPEN, bool MIX, bool TABLE,
          bool LAT>
__global__ void __launch_bounds__(M3_THREADS) __attribute__((amdgpu_waves_per_eu(WPE)))
k_m3_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, BinTable bt,
          const PosT* __restrict__ xp, const ValT* __restrict__ field, FieldGeom fg, ValT* __restrict__ up, int lda,
          int c0, int tz, int* __restrict__ lat_err) {
    static_assert(!(TABLE && LAT), "one binning form");
    constexpr int B = M3Geom<KS>::B, M3_TX = M3Geom<KS>::TX, M3_PL = M3Geom<KS>::PL;
    constexpr int RP = KS;                   // ring planes
    __shared__ ValT ring[RP * LDA * M3_PL];  // [plane][c][ry][nx]
    const int tid = threadIdx.x;
    const int ex = tid & 15, row = tid >> 4;
    // tiles cover the cells: N per periodic axis, N+1 (0..N) per open one
    const int Nx = m.nbin[0], Ny = m.nbin[1], Nz = m.nbin[2];
    const int ntx = (Nx + M3_TX - 1) / M3_TX, nty = (Ny + M3_TY - 1) / M3_TY;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int tx0 = (int)(tile % ntx) * M3_TX, ty0 = (int)((tile / ntx) % nty) * M3_TY;
    const int z0 = (int)(tile / ((int64_t)ntx * nty)) * tz;
    const int z1 = min(z0 + tz, Nz);
    const int nxo = min(M3_TX, Nx - tx0), nyo = min(M3_TY, Ny - ty0);
    const int64_t plane = (int64_t)Nx * Ny;

    // Ring planes move global -> registers -> LDS so that the global reads of
    // plane cz+3 are in flight while plane cz is computed.  A thread stages
    // whole nodes (all LDA components, adjacent in the field): the stores of
    // one component then write consecutive nodes, free of bank conflicts.
    constexpr int NN = (M3_PL + M3_THREADS - 1) / M3_THREADS;   // nodes per thread and plane
    ValT stage[NN][LDA];
    // Byte offset within a field plane of each node this thread stages (the
    // same in every plane; a plane fits in 2^31 bytes, checked at launch).
    // MIX: bit 31 marks a node whose x or y is not a real node of an open
    // axis, which every plane reads as stored.
    unsigned soff[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        const int node = tid + j * M3_THREADS;
        const int nx = node & 15, ry = node >> 4;
        const int gx = stored_node<OPEN>(m, fg, 0, tx0 - B + nx) - fg.lo[0];
        const int gy = stored_node<OPEN>(m, fg, 1, ty0 - B + ry) - fg.lo[1];
        soff[j] = (unsigned)(((gy * fg.ext[0] + gx) * lda + c0) * (int)sizeof(ValT));
        if (MIX && !(real_on_open_axis(m, 0, tx0 - B + nx) && real_on_open_axis(m, 1, ty0 - B + ry)))
            soff[j] |= 0x80000000u;
    }
    // MIX: byte offset of node `node` of a plane read as stored
    auto raw_off = [&](int node) -> unsigned {
        const int nx = node & 15, ry = node >> 4;
        const int gx = raw_node<OPEN>(m, fg, 0, tx0 - B + nx) - fg.lo[0];
        const int gy = raw_node<OPEN>(m, fg, 1, ty0 - B + ry) - fg.lo[1];
        return (unsigned)(((gy * fg.ext[0] + gx) * lda + c0) * (int)sizeof(ValT));
    };
    auto fetch_plane = [&](int n) {
        const int nz = stored_node<OPEN>(m, fg, 2, n) - fg.lo[2];
        const char* base = (const char*)(field + (int64_t)nz * fg.ext[0] * fg.ext[1] * lda);
        const char* braw = base;
        bool rz = true;
        if (MIX) {
            rz = real_on_open_axis(m, 2, n);
            const int nzr = raw_node<OPEN>(m, fg, 2, n) - fg.lo[2];
            braw = (const char*)(field + (int64_t)nzr * fg.ext[0] * fg.ext[1] * lda);
        }
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            const int node = tid + j * M3_THREADS;
            if (node < M3_PL) {
                const ValT* src;
                if (MIX && !(rz && (soff[j] >> 31) == 0u)) src = (const ValT*)(braw + raw_off(node));
                else src = (const ValT*)(base + soff[j]);
#pragma unroll
                for (int c = 0; c < LDA; ++c) stage[j][c] = src[c];
            }
        }
    };
    auto store_plane = [&](int n) {
        ValT* dst = ring + (n & (RP - 1)) * (LDA * M3_PL);
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            const int node = tid + j * M3_THREADS;
            if (node < M3_PL) {
#pragma unroll
                for (int c = 0; c < LDA; ++c) dst[c * M3_PL + node] = stage[j][c];
            }
        }
    };
    for (int n = z0 - B; n < z0 - B + KS; ++n) {
        fetch_plane(n);
        store_plane(n);
    }

    const bool plane_lane = ex >= B && ex - B < nxo && row < nyo;   // lanes that carry particles
    const int cx = tx0 + ex - B, cy = ty0 + row;
    const int key0 = (int)((int64_t)cy * Nx + cx);
    // Cell heads run two planes ahead of the computation, the permutation
    // entry of a plane's first particle one plane ahead, and its position is
    // loaded at the end of the plane before.
    // (f, n) is the cell's range [f, n) of the permutation, or (TABLE) its
    // key and count word, with t its first particle.
    auto heads = [&](int z, int& f, int& n, int& t) {
        f = 0;
        n = 0;
        t = 0;
        if (plane_lane && z < z1) {
            const int key = (int)((int64_t)z * plane) + key0;
            if constexpr (LAT) {
                f = key;
                n = 1;
            } else if constexpr (TABLE) {
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
        if constexpr (LAT) return n;
        else if constexpr (TABLE) return (int)(((unsigned)n >> count_shift(f)) & 255u);
        else return n - f;
    };
    // the cell's first particle, and its particle r < count
    auto lookup = [&](int f, int n, int t) {
        if constexpr (LAT) return f;
        else if constexpr (TABLE) return count_of(f, n) > 0 ? t : 0;
        else return n > f ? (perm ? perm[f] : f) : 0;
    };
    auto entry = [&](int f, int r) {
        if constexpr (LAT) return f + r;
        else if constexpr (TABLE) return bt.table[(int64_t)r * bt.ncells + f];
        else return perm ? perm[f + r] : f + r;
    };
    auto position = [&](int p, int n, PosT (&x)[3]) {
        x[0] = x[1] = x[2] = PosT(0);
        if (n > 0) {
            x[0] = xp[(int64_t)p * 3];
            x[1] = xp[(int64_t)p * 3 + 1];
            x[2] = xp[(int64_t)p * 3 + 2];
        }
    };
    int first_n, cnt_n, t_n, first_nn, cnt_nn, t_nn;
    heads(z0, first_n, cnt_n, t_n);
    heads(z0 + 1, first_nn, cnt_nn, t_nn);
    // Round 0's particle of a plane is loaded at the end of the plane before,
    // while the workgroup refills the ring.
    int pn = lookup(first_n, cnt_n, t_n);
    PosT xn[3];
    position(pn, count_of(first_n, cnt_n), xn);
    __syncthreads();

    bool lat_bad = false;   // (LAT) a particle found outside the cell it was taken for
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
        const int next = cz - B + KS;            // the plane that replaces plane cz - B
        if (more) fetch_plane(next);
        // one round: the lanes of cells with a particle left interpolate it
        auto round = [&](int p, bool has, const PosT (&x)[3]) {
            AccT wx[KS], wy[KS], wz[KS];
#pragma unroll
            for (int o = 0; o < KS; ++o) wx[o] = wy[o] = wz[o] = AccT(0);
            if (has) {
                int ci[3];
                tile_weights<KS, AccT>((AccT)cell_offset<PosT>(x[0], (PosT)m.minv[0], (PosT)m.dxi[0], ci[0]), wx);
                tile_weights<KS, AccT>((AccT)cell_offset<PosT>(x[1], (PosT)m.minv[1], (PosT)m.dxi[1], ci[1]), wy);
                tile_weights<KS, AccT>((AccT)cell_offset<PosT>(x[2], (PosT)m.minv[2], (PosT)m.dxi[2], ci[2]), wz);
                if constexpr (LAT) {
                    // The particle taken for this lane's cell must lie in it: the
                    // cell key's test on a periodic mesh (inside [min, max], INT
                    // truncation, index N wrapping to 0), from the cell indices
                    // the weights took.  No branch here: the outcome is kept and
                    // stored once, at the end of the kernel (a store amid the
                    // rounds kept the ring reads from being scheduled beside
                    // the weights).
                    const int cell[3] = {cx, cy, cz};
                    bool in = true;
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        in = in && x[d] >= (PosT)m.minv[d] && x[d] <= (PosT)m.maxv[d] &&
                             (ci[d] == cell[d] || (ci[d] == m.N[d] && cell[d] == 0));
                    lat_bad = lat_bad || !in;
                }
            }
            // (y,z) weights of the particle whose stencil takes this lane's column
            // as x offset k: that of lane ex + B - k (M'4: ex+1, ex, ex-1, ex-2)
            AccT a[KS][KS], b[KS][KS];
#pragma unroll
            for (int o = 0; o < KS; ++o) {
                a[0][o] = lane_from<B>(wy[o]);     b[0][o] = lane_from<B>(wz[o]);
                a[1][o] = lane_from<B - 1>(wy[o]); b[1][o] = lane_from<B - 1>(wz[o]);
                if constexpr (KS == 4) {
                    a[2][o] = lane_from<B - 2>(wy[o]); b[2][o] = lane_from<B - 2>(wz[o]);
                    a[3][o] = lane_from<B - 3>(wy[o]); b[3][o] = lane_from<B - 3>(wz[o]);
                }
            }
            // per component: column partial sums Ck = sum_oz b_k[oz] * sum_oy a_k[oy] f(oz, oy)
            // (one 4-node y slice of the column live at a time), then the particle's sum;
            // the components are stored together (one wide store per lane)
            ValT res[LDA];
#pragma unroll
            for (int c = 0; c < LDA; ++c) {
                AccT C[KS];
#pragma unroll
                for (int k = 0; k < KS; ++k) C[k] = AccT(0);
#pragma unroll
                for (int oz = 0; oz < KS; ++oz) {
                    const ValT* f = ring + ((cz - B + oz) & (RP - 1)) * (LDA * M3_PL) + c * M3_PL + row * 16 + ex;
                    AccT fv[KS];
#pragma unroll
                    for (int oy = 0; oy < KS; ++oy) fv[oy] = (AccT)f[16 * oy];
#pragma unroll
                    for (int k = 0; k < KS; ++k) {
                        AccT t = a[k][0] * fv[0];
#pragma unroll
                        for (int oy = 1; oy < KS; ++oy) t += a[k][oy] * fv[oy];
                        C[k] += b[k][oz] * t;
                    }
                }
                // x offset k of this lane's particle is the column of lane ex - B + k
                AccT u = wx[0] * lane_from<-B>(C[0]);
                u += wx[1] * lane_from<1 - B>(C[1]);
                if constexpr (KS == 4) {
                    u += wx[2] * lane_from<2 - B>(C[2]);
                    u += wx[3] * lane_from<3 - B>(C[3]);
                }
                res[c] = (ValT)u;
                // one component at a time: keeps the column reads of the next
                // component after this one's sums (register pressure)
                asm volatile("" : "+v"(res[c])::"memory");
            }
            if (has) {
#pragma unroll
                for (int c = 0; c < LDA; ++c) up[(int64_t)p * lda + c0 + c] = res[c];
            }
        };
        if (__ballot(cnt > 0) != 0ull) round(pn, cnt > 0, xn);
        for (int r = 1; __ballot(cnt > r) != 0ull; ++r) {
            PosT x[3];
            const int p = cnt > r ? entry(first, r) : 0;
            position(p, cnt - r, x);
            round(p, cnt > r, x);
        }
        pn = pnext;
        position(pn, count_of(first_n, cnt_n), xn);   // next plane's round 0
        __syncthreads();   // every row is done with plane cz-1: its slot takes plane cz+3
        if (more) store_plane(next);
        __syncthreads();
    }
    if constexpr (LAT) {
        if (__ballot(lat_bad) != 0ull && (threadIdx.x & 63) == 0) *lat_err = 1;
    }
}

// Per periodic axis the compact layout or a stored box holding the whole
// real range 0..N (read with the ghost_get node map); per open axis a box
// holding every node a stencil reaches, -1..N+2 (ghost size 2 or more).
// Beside an open axis a periodic axis stored as a box must hold -1..N+2 too:
// there ghost_get maps a node only when its open coordinates are real nodes,
// and the others are read as stored.
bool m2p3d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg) {
    if (m.dim != 3 || (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)) return false;
    // nodes a stencil reaches on an open axis (cells 0..N): -1..N+2 (M'4), 0..N+1 (BSP2)
    const int below = kernel == PPMG_KERNEL_MP4 ? 1 : 0, above = kernel == PPMG_KERNEL_MP4 ? 2 : 1;
    bool open = false, swapped = false;
    for (int d = 0; d < 3; ++d) {
        open = open || !m.periodic[d];
        swapped = swapped || fg.swap[d];
    }
    // there a periodic box must hold every node a stencil reaches as stored
    if (open && swapped)
        for (int d = 0; d < 3; ++d)
            if (fg.swap[d] && (fg.lo[d] > -below || fg.lo[d] + fg.ext[d] - 1 < m.N[d] + above)) return false;
    for (int d = 0; d < 3; ++d) {
        if (m.N[d] < 4) return false;
        const int hi = fg.lo[d] + fg.ext[d] - 1;
        if (m.periodic[d]) {
            const bool compact = fg.lo[d] == 0 && fg.ext[d] == m.N[d];
            const bool boxed = fg.lo[d] <= 0 && hi >= m.N[d];
            if (!compact && !boxed) return false;
        } else if (fg.lo[d] > -below || hi < m.N[d] + above) {
            return false;
        }
    }
    // the ring fetch addresses a stored plane with 32-bit byte offsets
    return (int64_t)fg.ext[0] * fg.ext[1] * fg.lda * 8 < ((int64_t)1 << 31);
}

// The gather reads a slot-table binning on a periodic mesh (whatever the
// field's layout).
bool m2p3d_table_applicable(const DevMesh& m) { return m.dim == 3 && m.periodic[0] && m.periodic[1] && m.periodic[2]; }

// One launch of the gather for components c0 .. c0+LL-1; the z chunk comes
// from the resident workgroup count (pick_z_chunk_balanced: each chunk loads
// 3 extra node planes, cheap beside the balance short chunks give).
template <int KS, int LL, typename PosT, typename ValT, typename AccT, bool OPEN, bool MIX, bool TABLE = false,
          bool LAT = false>
static void launch_m3_open(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const FieldGeom& fg, int c0,
                           const ValT* field, ValT* up, hipStream_t s) {
    constexpr int TX = M3Geom<KS>::TX;
    constexpr int WPE = KS == 4 && (LL == 4 || MIX) ? 3 : 4;
    auto kernel = k_m3_tile<KS, LL, PosT, ValT, AccT, WPE, OPEN, MIX, TABLE, LAT>;
    static std::atomic<int> slots{0};
    if (slots.load() == 0) slots.store(resident_workgroups(kernel, M3_THREADS));
    const int tx = (m.nbin[0] + TX - 1) / TX, ty = (m.nbin[1] + M3_TY - 1) / M3_TY;
    const int tz = pick_z_chunk_balanced((int64_t)tx * ty, m.nbin[2], slots.load(), 32);
    dim3 grid((unsigned)((int64_t)tx * ty * ((m.nbin[2] + tz - 1) / tz)));
    const BinTable bt{plan->count.as<uint32_t>(), plan->table.as<int>(), plan->ncells};
    hipLaunchKernelGGL(kernel, grid, dim3(M3_THREADS), 0, s, m, plan->start.as<int>(),
                       tile_perm(plan), bt, xp, field, fg, up, fg.lda, c0, tz, plan->err.as<int>() + 6);
}

// OPEN: some axis is open (a fully periodic mesh compiles without the clamp);
// MIX: and some periodic axis is stored as a box.
template <int KS, int LL, typename PosT, typename ValT, typename AccT>
static void launch_m3(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const FieldGeom& fg, int c0,
                      const ValT* field, ValT* up, hipStream_t s) {
    // a slot table or an unchecked lattice is binned on periodic meshes only (m2p3d_fast)
    if (plan->bin_lattice)
        launch_m3_open<KS, LL, PosT, ValT, AccT, false, false, false, true>(plan, m, xp, fg, c0, field, up, s);
    else if (plan->bin_table)
        launch_m3_open<KS, LL, PosT, ValT, AccT, false, false, true>(plan, m, xp, fg, c0, field, up, s);
    else if (m.periodic[0] && m.periodic[1] && m.periodic[2])
        launch_m3_open<KS, LL, PosT, ValT, AccT, false, false>(plan, m, xp, fg, c0, field, up, s);
    else if (fg.swap[0] || fg.swap[1] || fg.swap[2])
        launch_m3_open<KS, LL, PosT, ValT, AccT, true, true>(plan, m, xp, fg, c0, field, up, s);
    else
        launch_m3_open<KS, LL, PosT, ValT, AccT, true, false>(plan, m, xp, fg, c0, field, up, s);
}

template <int KS, typename PosT, typename ValT, typename AccT>
static int m2p3d_typed(ppmg_plan* plan, const DevMesh& m, int64_t np, const PosT* xp, const FieldGeom& fg,
                       const ValT* field, ValT* up, hipStream_t s) {
    const int lda = fg.lda;
    (void)np;
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL) launch_m3<KS, LL, PosT, ValT, AccT>(plan, m, xp, fg, c0, field, up, s)
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
// when some particle lies outside the sub, so the caller can use the general
// path, which reads such particles' stencils with bounds checks.
//
// Arithmetic: FP64 in every form.  On gfx950 (packed FP32 off) an FP64 FMA
// issues at the FP32 rate, so the FP32-compensated form keeps FP32 only where
// it saves bytes (field in HBM and LDS, output) and sums its 64 products in
// FP64 registers -- more accurate than compensated FP32 at no extra cost.
//
// A plan whose last binning found a lattice in cell order (particle p in
// cell p, one per cell: a fresh lattice of remeshed particles) takes the
// next one on the same mesh as that lattice without a binning pass: the
// gather checks every particle's cell key and, should one differ, the call
// bins and gathers again.
int m2p3d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp,
               const FieldGeom& fg, const void* field, void* up, bool prebinned, hipStream_t s) {
    const bool periodic = m2p3d_table_applicable(m);
    int64_t nbad = prebinned ? plan->prebin_nbad : 0;
    if (!prebinned && bin_particles(plan, m, prec, np, xp, s, &nbad, periodic, periodic)) return PPMG_ERROR_FATAL;
    if (nbad > 0) return 1;
    // (a prebinning of a mesh the gather cannot read as a table or check as a lattice)
    if ((plan->bin_table || plan->bin_lattice) && !periodic && materialize_heads(plan, s)) return PPMG_ERROR_FATAL;
    mark(plan, s);
    if (np <= 0) return 0;
    auto gather = [&]() -> int {
#define TYPED(KS)                                                                                              \
    do {                                                                                                       \
        if (prec == PPMG_PREC_F64)                                                                             \
            return m2p3d_typed<KS, double, double, double>(plan, m, np, (const double*)xp, fg,                 \
                                                           (const double*)field, (double*)up, s);              \
        if (prec == PPMG_PREC_F32C)                                                                            \
            return m2p3d_typed<KS, double, float, double>(plan, m, np, (const double*)xp, fg,                  \
                                                          (const float*)field, (float*)up, s);                 \
        return m2p3d_typed<KS, float, float, double>(plan, m, np, (const float*)xp, fg, (const float*)field,   \
                                                     (float*)up, s);                                           \
    } while (0)
        if (kernel == PPMG_KERNEL_BSP2) TYPED(2);
        TYPED(4);
#undef TYPED
    };
    if (!plan->bin_lattice) return gather();
    int* lat_err = plan->err.as<int>() + 6;
    if (fill_zero(lat_err, sizeof(int), s)) return PPMG_ERROR_FATAL;
    if (gather()) return PPMG_ERROR_FATAL;
    int off = 0;
    PPMG_HIP(hipMemcpyAsync(&off, lat_err, sizeof(int), hipMemcpyDeviceToHost, s));
    PPMG_HIP(hipStreamSynchronize(s));
    if (!off) {
        plan->lattice_checked = true;   // the plan's binning is the lattice (heads written on demand)
        return 0;
    }
    // not that lattice: bin and gather again (and bin the next 16 inputs)
    plan->lattice_pred = false;
    plan->lattice_skip = 16;
    if (bin_particles(plan, m, prec, np, xp, s, &nbad, periodic, false)) return PPMG_ERROR_FATAL;
    if (nbad > 0) return 1;
    return gather();
}

}  // namespace ppmg
