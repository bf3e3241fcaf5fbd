// ppmg_tilebox.h -- the stored nodes of a field box besides the ones the
// tiled deposits write (2D and 3D, M'4 and BSP2): node N of periodic axes
// whose stored box holds the whole real range, and the ghost layers.  Shared
// by ppmg_p2m2d.hip and ppmg_p2m3d.hip.
#pragma once
#include "ppmg_internal.h"
#include "ppmg_sort.h"

namespace ppmg {

// Weight tables of ppm_rmsh_remesh (wx1/wx2/wx3, four per particle and axis),
// used in place of the positions' weights (3D M'4 only).
template <typename ValT> struct WTab {
    const ValT* w[3];
};

// Stored nodes besides the unique ones, for axes whose stored box holds the
// whole real range 0..N.  Per axis: R = stored real nodes (N compact, N+1
// boxed), O = ext - R stored nodes outside 0..N.

// Real nodes with index N on a boxed periodic axis (n.d = N; -1 on other
// axes): the fold leaves node N equal to node 0, so they take the value the
// tiles wrote there.  Enumerated as the face i = N, then j = N (i < N), then
// k = N (i, j < N), over the real nodes (r per axis, b = 1 where node N is a
// copy).
template <typename ValT>
__global__ void __launch_bounds__(256) k_box_dup(FieldGeom fg, int3 n, int3 r, int3 b, ValT* __restrict__ field,
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
// with weight tables (wt.w[0] set) the weights are the tables'.  KS: stencil
// width (4 M'4, 2 BSP2); DIM 2: one plane (no z stencil), positions of two
// coordinates.  TABLE: the binning is a slot table instead of heads and a
// permutation; a cell's particles come in the same order.
template <int DIM, int KS, typename PosT, typename ValT, int LL, bool TABLE>
__global__ void __launch_bounds__(256) k_box_shell(DevMesh m, FieldGeom fg, int3 r, const int* __restrict__ start,
                                                  const int* __restrict__ perm, BinTable bt,
                                                  const PosT* __restrict__ xp, const ValT* __restrict__ up, int c0,
                                                  ValT* __restrict__ field, int64_t nshell, WTab<ValT> wt) {
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
    constexpr int B = Stencil<KS>::B;
    // The stencil offsets o of each axis whose cell n + B - o exists, walked
    // in the order z, y, x (x fastest) with the next cell's list head (count
    // and first particle) fetched while the current cell is summed: the same
    // additions in the same order as a plain loop over the KS^DIM offsets,
    // with one cell's lookups in flight instead of none.
    int olo[3], ohi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        olo[d] = d < DIM ? max(0, n[d] + B - (m.nbin[d] - 1)) : 0;
        ohi[d] = d < DIM ? min(KS - 1, n[d] + B) : 0;
    }
    auto key_of = [&](int ox, int oy, int oz) {
        const int cz = DIM == 3 ? n[2] + B - oz : 0;
        return (n[0] + B - ox) + m.nbin[0] * ((n[1] + B - oy) + m.nbin[1] * cz);
    };
    // list of cell `key`: particles q in [b, e), particle p0 at q = b (if any)
    auto head = [&](int key, int& b, int& e, int& p0) {
        b = TABLE ? 0 : start[key];
        e = TABLE ? (int)((bt.count[count_word(key)] >> count_shift(key)) & 255u) : start[key + 1];
        p0 = e > b ? (TABLE ? bt.table[key] : perm ? perm[b] : b) : 0;
    };
    auto particle_of = [&](int key, int q) {
        return TABLE ? bt.table[(int64_t)q * bt.ncells + key] : perm ? perm[q] : q;
    };
    // one particle's contribution from its position and strengths
    auto add_vals = [&](const PosT (&x)[DIM], const ValT (&u)[LL], int ox, int oy, int oz) {
        ValT fr[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            int ci;
            fr[d] = (ValT)cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], ci);
        }
        ValT w = tile_weight_at<KS, ValT>(fr[0], ox) * tile_weight_at<KS, ValT>(fr[1], oy);
        if constexpr (DIM == 3) w = w * tile_weight_at<KS, ValT>(fr[2], oz);
#pragma unroll
        for (int c = 0; c < LL; ++c) acc[c] += (double)(w * u[c]);
    };
    auto add = [&](int p, int ox, int oy, int oz) {
        if (DIM == 3 && KS == 4 && wt.w[0]) {
            const ValT w =
                (wt.w[0][(int64_t)p * 4 + ox] * wt.w[1][(int64_t)p * 4 + oy]) * wt.w[2][(int64_t)p * 4 + oz];
#pragma unroll
            for (int c = 0; c < LL; ++c) acc[c] += (double)(w * up[(int64_t)p * fg.lda + c0 + c]);
            return;
        }
        PosT x[DIM];
        ValT u[LL];
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[d] = xp[(int64_t)p * DIM + d];
#pragma unroll
        for (int c = 0; c < LL; ++c) u[c] = up[(int64_t)p * fg.lda + c0 + c];
        add_vals(x, u, ox, oy, oz);
    };
    // A node outside the real range on an axis f of a periodic 3D M'4 box
    // meets one stencil offset there (node -1: cell 0, node N+1: cell N-1),
    // so its cells are at most 4 x 4 along the other two axes: they are taken
    // as 16 fixed slots (outer free axis, then inner), all 16 list heads
    // loaded at once and the first particles (position or table weights, and
    // strengths) of 4 slots at a time, then summed in the same order as the
    // walk below.
    int f = -1;
    if constexpr (DIM == 3 && KS == 4) {
        if (olo[0] <= ohi[0] && olo[1] <= ohi[1] && olo[2] <= ohi[2]) {
#pragma unroll
            for (int d = 2; d >= 0; --d)
                if (olo[d] == ohi[d]) f = d;
        }
    }
    if (f >= 0) {
        const int A = f == 2 ? 1 : 2, Bx = f == 0 ? 1 : 0;   // outer and inner free axes
        auto offs = [&](int sl, int& ox, int& oy, int& oz) {
            int o[3];
            o[f] = olo[f];
            o[A] = olo[A] + (sl >> 2);
            o[Bx] = olo[Bx] + (sl & 3);
            ox = o[0];
            oy = o[1];
            oz = o[2];
            return o[A] <= ohi[A] && o[Bx] <= ohi[Bx];
        };
        int cb[16], ce[16], cp[16];
#pragma unroll
        for (int sl = 0; sl < 16; ++sl) {
            int ox, oy, oz;
            cb[sl] = ce[sl] = cp[sl] = 0;
            if (offs(sl, ox, oy, oz)) head(key_of(ox, oy, oz), cb[sl], ce[sl], cp[sl]);
        }
#pragma unroll
        for (int s0 = 0; s0 < 16; s0 += 4) {
            // one batch's particle loads in flight at a time (register bound)
            asm volatile("" ::: "memory");
            // per slot: the position, or with weight tables the particle's
            // three table weights at the slot's offsets (held in the same
            // registers: a weight converts to the position type and back
            // exactly)
            PosT xs[4][DIM];
            ValT us[4][LL];
            const bool tab = wt.w[0] != nullptr;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (ce[s0 + j] > cb[s0 + j]) {
                    const int p = cp[s0 + j];
                    if (tab) {
                        int ox, oy, oz;
                        offs(s0 + j, ox, oy, oz);
                        xs[j][0] = (PosT)wt.w[0][(int64_t)p * 4 + ox];
                        xs[j][1] = (PosT)wt.w[1][(int64_t)p * 4 + oy];
                        xs[j][2] = (PosT)wt.w[2][(int64_t)p * 4 + oz];
                    } else {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) xs[j][d] = xp[(int64_t)p * DIM + d];
                    }
#pragma unroll
                    for (int c = 0; c < LL; ++c) us[j][c] = up[(int64_t)p * fg.lda + c0 + c];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int sl = s0 + j;
                if (ce[sl] > cb[sl]) {
                    int ox, oy, oz;
                    offs(sl, ox, oy, oz);
                    if (tab) {
                        const ValT w = ((ValT)xs[j][0] * (ValT)xs[j][1]) * (ValT)xs[j][2];
#pragma unroll
                        for (int c = 0; c < LL; ++c) acc[c] += (double)(w * us[j][c]);
                    } else {
                        add_vals(xs[j], us[j], ox, oy, oz);
                    }
                    for (int q = cb[sl] + 1; q < ce[sl]; ++q) add(particle_of(key_of(ox, oy, oz), q), ox, oy, oz);
                }
            }
        }
    } else if (olo[0] <= ohi[0] && olo[1] <= ohi[1] && olo[2] <= ohi[2]) {
        int ox = olo[0], oy = olo[1], oz = olo[2];
        int key = key_of(ox, oy, oz), b, e, p0;
        head(key, b, e, p0);
        for (;;) {
            // the cell after this one
            int nx = ox + 1, ny = oy, nz = oz;
            if (nx > ohi[0]) {
                nx = olo[0];
                if (++ny > ohi[1]) {
                    ny = olo[1];
                    ++nz;
                }
            }
            const bool more = nz <= ohi[2];
            int nkey = 0, nb = 0, ne = 0, np0 = 0;
            if (more) {
                nkey = key_of(nx, ny, nz);
                head(nkey, nb, ne, np0);
            }
            if (e > b) {
                add(p0, ox, oy, oz);
                for (int q = b + 1; q < e; ++q) add(particle_of(key, q), ox, oy, oz);
            }
            if (!more) break;
            ox = nx;
            oy = ny;
            oz = nz;
            key = nkey;
            b = nb;
            e = ne;
            p0 = np0;
        }
    }
    ValT* out = field + (((int64_t)(n[2] - fg.lo[2]) * e1 + (n[1] - fg.lo[1])) * e0 + (n[0] - fg.lo[0])) * fg.lda + c0;
#pragma unroll
    for (int c = 0; c < LL; ++c) out[c] = (ValT)acc[c];
}

// Fills the stored nodes of `field` that the tiles do not write (see above),
// after the tile launches of a deposit.
template <int DIM, int KS, typename PosT, typename ValT>
int fill_stored_box(ppmg_plan* plan, const DevMesh& m, const PosT* xp, const ValT* up, const FieldGeom& fg,
                    ValT* field, const WTab<ValT>& wt, hipStream_t s) {
    const int lda = fg.lda;
    int R[3], B[3], D[3];
    for (int d = 0; d < 3; ++d) {
        const bool compact = m.periodic[d] && fg.lo[d] == 0 && fg.ext[d] == m.N[d];
        B[d] = d < DIM && m.periodic[d] && !compact;
        R[d] = d >= DIM ? 1 : compact ? m.N[d] : m.N[d] + 1;
        D[d] = B[d] ? m.N[d] : -1;
    }
    const int3 n3 = make_int3(D[0], D[1], D[2]), r3 = make_int3(R[0], R[1], R[2]),
               b3 = make_int3(B[0], B[1], B[2]);
    const int64_t ndup = (int64_t)R[0] * R[1] * R[2] - (int64_t)(R[0] - B[0]) * (R[1] - B[1]) * (R[2] - B[2]);
    if (ndup > 0)
        hipLaunchKernelGGL(k_box_dup<ValT>, dim3((unsigned)((ndup + 255) / 256)), dim3(256), 0, s, fg, n3, r3, b3,
                           field, ndup);
    const int64_t nshell = (int64_t)fg.ext[0] * fg.ext[1] * fg.ext[2] - (int64_t)R[0] * R[1] * R[2];
    if (nshell > 0) {
        const unsigned g = (unsigned)((nshell + 255) / 256);
        const BinTable bt{plan->count.as<uint32_t>(), plan->table.as<int>(), plan->ncells};
        const bool table = plan->bin_table;
        if (!table && materialize_heads(plan, s)) return 1;
        for (int c0 = 0; c0 < lda; c0 += 4) {
            const int L = lda - c0 < 4 ? lda - c0 : 4;
#define SH(LL, TB) hipLaunchKernelGGL((k_box_shell<DIM, KS, PosT, ValT, LL, TB>), dim3(g), dim3(256), 0, s, m, fg, r3, \
                                      plan->start.as<int>(), tile_perm(plan), bt, xp, up, c0, field, nshell, wt)
#define SHL(LL)                      \
    do {                             \
        if (table) SH(LL, true);     \
        else SH(LL, false);          \
    } while (0)
            if (L == 1) SHL(1);
            else if (L == 2) SHL(2);
            else if (L == 3) SHL(3);
            else SHL(4);
#undef SHL
#undef SH
        }
    }
    PPMG_HIP(hipGetLastError());
    return 0;
}

}  // namespace ppmg
