// ppmg_p2m3d.hip -- 3D M'4 P2M fast path for gfx950: tiled z-march deposit.
//
// Replaces the scatter loop of ppm_interp_p2m (ppm_interp_p2m.f:783-1377,
// 2540-3055) plus its periodic ghost fold (:3670-3716) for fully periodic
// meshes written in the compact layout (unique nodes only).
//
// Work decomposition (one wavefront per workgroup):
//   * a workgroup owns a 13 x 13 node tile in x,y and 32 node planes in z;
//   * it marches the cell planes cz = z0-2 .. z1 (the cells whose M'4 stencil
//     touches its planes) and keeps a ring of 4 node planes in LDS: cell plane
//     cz deposits into node planes cz-1 .. cz+2, after which node plane cz-1
//     is complete and is written out with coalesced row stores;
//   * the 16 x 16 cells whose stencils touch the tile (13 + 3 halo) are taken
//     64 at a time, one cell per lane; each lane walks its cell's particles
//     (binned, in stable order) and adds the owned part of the 4x4x4 stencil.
// Lanes of one instruction hold distinct cells, hence write distinct nodes,
// and one wavefront owns the ring: the order of every node's additions is a
// fixed function of the binning, so results are bitwise reproducible with no
// global float atomics.  The periodic fold happens through the index wrap of
// the cells around the tile.
#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

constexpr int P3_TT = 13;   // owned nodes per tile edge (x, y)
constexpr int P3_TE = 16;   // cells per tile edge incl. halo (TT + 3)
constexpr int P3_TZ = 32;   // owned node planes per z chunk
constexpr int P3_PL = P3_TT * P3_TT;

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

// Sorted particle records: fractional cell offsets (INT truncation as in
// ppm_interp_p2m.f:800-822, in the position precision) and strengths.
template <typename PosT, typename FracT, typename ValT>
__global__ void __launch_bounds__(256)
k_p3_sorted(const int* __restrict__ perm, const PosT* __restrict__ xp, const ValT* __restrict__ up, int lda,
            int64_t np, DevMesh m, FracT* __restrict__ fr, ValT* __restrict__ us) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    int p = perm[s];
    const PosT* x = xp + (int64_t)p * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int i;
        fr[s * 3 + d] = (FracT)cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], i);
    }
    for (int c = 0; c < lda; ++c) us[s * lda + c] = up[(int64_t)p * lda + c];
}

// Shift a value one to three lanes up inside each 16-lane row (DPP row_shr,
// bound_ctrl: lanes shifted in from outside the row read 0).
template <int K>
__device__ __forceinline__ double row_shr(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x110 + K, 0xF, 0xF, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x110 + K, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <int K>
__device__ __forceinline__ float row_shr(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + K, 0xF, 0xF, true));
}

// Tiled deposit.  Lane = one cell of the 16 x 16 window: ex = lane & 15 along
// x, so each 16-lane DPP row is a run of 16 consecutive cells at fixed y.  In
// round r every lane forms the contributions of the r-th particle of its cell
// (zero when it has none); the four x-offsets that land on one node come from
// four neighbouring lanes, and are summed across the row with DPP shifts:
//   node(ex - 3) = v0(ex) + v1(ex-1) + v2(ex-2) + v3(ex-3)
// so each lane adds ONE value per (y, z offset, component) to the ring -- 48
// LDS atomics per particle for lda = 3 instead of 192.  The ring is stored
// component-major (SoA) so a row's 16 lanes hit consecutive 8-byte words.
template <int LDA, typename ValT, typename FracT>
__global__ void __launch_bounds__(64)
k_p3_tile(DevMesh m, const int* __restrict__ start, const FracT* __restrict__ fr, const ValT* __restrict__ us,
          int lda, int c0, ValT* __restrict__ field) {
    using A = LdsAcc<ValT>;
    using slot_t = typename A::slot_t;
    constexpr int PLANE = P3_PL * LDA;            // slots per ring plane: [c][node]
    __shared__ slot_t ring[4 * PLANE];
    const int lane = threadIdx.x;
    const int Nx = m.N[0], Ny = m.N[1], Nz = m.N[2];
    const int tx0 = blockIdx.x * P3_TT, ty0 = blockIdx.y * P3_TT, z0 = blockIdx.z * P3_TZ;
    const int z1 = min(z0 + P3_TZ, Nz);
    const int nxo = min(P3_TT, Nx - tx0), nyo = min(P3_TT, Ny - ty0);
    for (int i = lane; i < 4 * PLANE; i += 64) A::zero(ring[i]);

    const int ex = lane & 15, erow = lane >> 4;
    const int lx = ex - 3;                          // node this lane owns after the row sum
    const bool xown = lx >= 0 && lx < nxo;
    int cx = tx0 - 2 + ex;
    cx %= Nx; if (cx < 0) cx += Nx;
    int keyxy[4], ey4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int ey = q * 4 + erow;
        ey4[q] = ey;
        int cy = ty0 - 2 + ey;
        cy %= Ny; if (cy < 0) cy += Ny;
        keyxy[q] = cx + Nx * cy;
    }
    __syncthreads();
    const int64_t plane_cells = (int64_t)Nx * Ny;

    for (int cz = z0 - 2; cz <= z1; ++cz) {
        int czw = cz % Nz;
        if (czw < 0) czw += Nz;
        const int64_t kz = plane_cells * czw;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int key = (int)(kz + keyxy[q]);
            const int first = start[key];
            const int cnt = start[key + 1] - first;
            const int ly0 = ey4[q] - 3;              // node row of stencil offset oy = 0
            for (int r = 0; __ballot(cnt > r) != 0ull; ++r) {
                const bool has = cnt > r;
                const int sp = has ? first + r : 0;
                ValT wx[4], wy[4], wz[4], u[LDA];
                mp4_unrolled<ValT>((ValT)fr[3 * (int64_t)sp], wx);
                mp4_unrolled<ValT>((ValT)fr[3 * (int64_t)sp + 1], wy);
                mp4_unrolled<ValT>((ValT)fr[3 * (int64_t)sp + 2], wz);
#pragma unroll
                for (int c = 0; c < LDA; ++c) u[c] = has ? us[(int64_t)sp * lda + c0 + c] : ValT(0);
#pragma unroll
                for (int oy = 0; oy < 4; ++oy) {
                    const int ly = ly0 + oy;
                    const bool own = xown && ly >= 0 && ly < nyo;
                    const int node = lx + P3_TT * ly;
#pragma unroll
                    for (int oz = 0; oz < 4; ++oz) {
                        slot_t* pl = ring + ((cz - 1 + oz) & 3) * PLANE;
                        const ValT wyz = wy[oy] * wz[oz];
#pragma unroll
                        for (int c = 0; c < LDA; ++c) {
                            const ValT t = wyz * u[c];
                            ValT v = wx[0] * t;
                            v += row_shr<1>(wx[1] * t);
                            v += row_shr<2>(wx[2] * t);
                            v += row_shr<3>(wx[3] * t);
                            if (own) A::add(pl + c * P3_PL + node, v);
                        }
                    }
                }
            }
        }
        __syncthreads();
        // node plane cz-1 has all its cells: write it if owned, then recycle the slot
        const int n = cz - 1;
        slot_t* pl = ring + (n & 3) * PLANE;
        if (n >= z0 && n < z1) {
            const int cntw = nyo * nxo * LDA;
            ValT* out = field + (int64_t)n * plane_cells * lda + c0;
            for (int i = lane; i < cntw; i += 64) {
                int c = i % LDA;
                int nd = i / LDA;
                int x = nd % nxo, y = nd / nxo;
                out[((int64_t)(ty0 + y) * Nx + (tx0 + x)) * lda + c] = (ValT)A::value(pl[c * P3_PL + x + P3_TT * y]);
            }
        }
        for (int i = lane; i < PLANE; i += 64) A::zero(pl[i]);
        __syncthreads();
    }
}

static inline unsigned nblk(int64_t n, int b) {
    int64_t g = (n + b - 1) / b;
    return (unsigned)(g > 0 ? g : 1);
}

bool p2m3d_fast_applicable(const DevMesh& m, int kernel, int prec, const FieldGeom& fg) {
    if (m.dim != 3 || kernel != PPMG_KERNEL_MP4) return false;
    for (int d = 0; d < 3; ++d) {
        if (!m.periodic[d] || m.N[d] < 4) return false;
        if (fg.lo[d] != 0 || fg.ext[d] != m.N[d]) return false;
    }
    (void)prec;
    return true;
}

template <typename PosT, typename FracT, typename ValT>
static int p2m3d_typed(ppmg_plan* plan, const DevMesh& m, int64_t np, const PosT* xp, int lda, const ValT* up,
                       ValT* field, hipStream_t s) {
    if (plan->wsorted.ensure(sizeof(FracT) * 3 * (size_t)(np > 0 ? np : 1)) ||
        plan->usorted.ensure(sizeof(ValT) * (size_t)lda * (size_t)(np > 0 ? np : 1)))
        return PPMG_ERROR_FATAL;
    FracT* fr = plan->wsorted.as<FracT>();
    ValT* us = plan->usorted.as<ValT>();
    if (np > 0) {
        hipLaunchKernelGGL((k_p3_sorted<PosT, FracT, ValT>), dim3(nblk(np, 256)), dim3(256), 0, s, plan->perm.as<int>(),
                           xp, up, lda, np, m, fr, us);
        PPMG_HIP(hipGetLastError());
    }
    mark(plan, s);
    dim3 grid((unsigned)((m.N[0] + P3_TT - 1) / P3_TT), (unsigned)((m.N[1] + P3_TT - 1) / P3_TT),
              (unsigned)((m.N[2] + P3_TZ - 1) / P3_TZ));
    const int* start = plan->start.as<int>();
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL) hipLaunchKernelGGL((k_p3_tile<LL, ValT, FracT>), grid, dim3(64), 0, s, m, start, fr, us, lda, c0, field)
        if (L == 1) TL(1);
        else if (L == 2) TL(2);
        else if (L == 3) TL(3);
        else TL(4);
#undef TL
        PPMG_HIP(hipGetLastError());
    }
    mark(plan, s);
    return 0;
}

int p2m3d_fast(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp, int lda, const void* up,
               void* field, hipStream_t s) {
    if (prec == PPMG_PREC_F64)
        return p2m3d_typed<double, double, double>(plan, m, np, (const double*)xp, lda, (const double*)up,
                                                   (double*)field, s);
    if (prec == PPMG_PREC_F32C)
        return p2m3d_typed<double, float, float>(plan, m, np, (const double*)xp, lda, (const float*)up, (float*)field,
                                                 s);
    return p2m3d_typed<float, float, float>(plan, m, np, (const float*)xp, lda, (const float*)up, (float*)field, s);
}

}  // namespace ppmg
