// ppmg_m2p3d.hip -- 3D M'4 M2P fast path for gfx950: tiled z-march gather.
//
// Replaces the gather loop of ppm_interp_m2p (ppm_interp_m2p.f:895-1240) for
// fully periodic meshes stored in the compact layout; the periodic ghost fill
// the reference gets from ppm_map_field_ghost (ghost_get) is the index wrap of
// the node planes loaded around the tile.
//
// A workgroup (4 wavefronts) owns the particles of a 13 x 13 cell tile in x,y
// over 32 cell planes in z.  The particles are binned first (stable counting
// sort), so the particles of one cell row of the tile are contiguous.  The
// 16 x 16 nodes those cells read (13 + 3) are kept for 4 node planes in an
// LDS ring: cell plane cz reads node planes cz-1 .. cz+2; afterwards plane
// cz-1 is replaced by plane cz+3, loaded with coalesced row reads.  Every node
// of the field is therefore read from HBM (16/13)^2 times, and every particle
// gathers its 4x4x4 stencil from LDS.  Output order is the caller's particle
// order (written through the permutation).
#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

constexpr int M3_TT = 13;                 // owned cells per tile edge (x, y)
constexpr int M3_TE = 16;                 // nodes per tile edge (TT + 3)
constexpr int M3_TZ = 32;                 // cell planes per z chunk
constexpr int M3_PITCH = 20;              // ring row pitch in nodes (16 used; padding spreads LDS banks)
constexpr int M3_PL = M3_PITCH * M3_TE;   // node slots per ring plane and component
constexpr int M3_THREADS = 256;

template <int LDA, typename PosT, typename ValT>
__global__ void __launch_bounds__(M3_THREADS)
k_m3_tile(DevMesh m, const int* __restrict__ start, const int* __restrict__ perm, const PosT* __restrict__ xp,
          const ValT* __restrict__ field, ValT* __restrict__ up, int lda, int c0) {
    using Acc = typename AccOf<ValT>::type;
    __shared__ ValT ring[4 * M3_PL * LDA];
    __shared__ int row_first[M3_TT];
    __shared__ int row_pref[M3_TT + 1];
    const int tid = threadIdx.x;
    const int Nx = m.N[0], Ny = m.N[1], Nz = m.N[2];
    const int tx0 = blockIdx.x * M3_TT, ty0 = blockIdx.y * M3_TT, z0 = blockIdx.z * M3_TZ;
    const int z1 = min(z0 + M3_TZ, Nz);
    const int nxo = min(M3_TT, Nx - tx0), nyo = min(M3_TT, Ny - ty0);
    const int64_t plane = (int64_t)Nx * Ny;

    auto load_plane = [&](int n) {
        ValT* dst = ring + (n & 3) * (M3_PL * LDA);
        int nz = n % Nz;
        if (nz < 0) nz += Nz;
        // coalesced in the field (component fastest), component-major in the ring
        for (int i = tid; i < M3_TE * M3_TE * LDA; i += M3_THREADS) {
            int c = i % LDA;
            int node = i / LDA;
            int nx = node & (M3_TE - 1), ny = node / M3_TE;
            int gx = tx0 - 1 + nx;
            int gy = ty0 - 1 + ny;
            gx %= Nx; if (gx < 0) gx += Nx;
            gy %= Ny; if (gy < 0) gy += Ny;
            dst[c * M3_PL + nx + M3_PITCH * ny] = field[((int64_t)nz * plane + (int64_t)gy * Nx + gx) * lda + c0 + c];
        }
    };
    for (int n = z0 - 1; n <= z0 + 2; ++n) load_plane(n);
    __syncthreads();

    for (int cz = z0; cz < z1; ++cz) {
        if (tid < nyo) {
            int key = (int)((int64_t)cz * plane + (int64_t)(ty0 + tid) * Nx + tx0);
            int b = start[key];
            row_first[tid] = b;
            row_pref[tid + 1] = start[key + nxo] - b;
        }
        __syncthreads();
        if (tid == 0) {
            row_pref[0] = 0;
            for (int r = 0; r < nyo; ++r) row_pref[r + 1] += row_pref[r];
        }
        __syncthreads();
        const int total = row_pref[nyo];
        for (int i = tid; i < total; i += M3_THREADS) {
            int r = 0;
            while (r + 1 < nyo && row_pref[r + 1] <= i) ++r;
            const int s = row_first[r] + (i - row_pref[r]);
            const int p = perm[s];
            const PosT* x = xp + (int64_t)p * 3;
            ValT a[3][4];
            int cell[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                int ci;
                PosT frac = cell_offset<PosT>(x[d], (PosT)m.minv[d], (PosT)m.dxi[d], ci);
                mp4_unrolled<ValT>((ValT)frac, a[d]);
                if (ci >= m.N[d]) ci -= m.N[d];
                cell[d] = ci;
            }
            const int lx = cell[0] - tx0, ly = cell[1] - ty0;   // in [0, 13): node lx + ox of the ring
            if ((unsigned)lx >= (unsigned)nxo || (unsigned)ly >= (unsigned)nyo) continue;   // not binned here
            // one partial sum per z offset: four independent FMA chains per component
            Acc acc[4][LDA];
#pragma unroll
            for (int oz = 0; oz < 4; ++oz)
#pragma unroll
                for (int c = 0; c < LDA; ++c) acc[oz][c].init();
#pragma unroll
            for (int ox = 0; ox < 4; ++ox) {
#pragma unroll
                for (int oy = 0; oy < 4; ++oy) {
                    const ValT wxy = a[0][ox] * a[1][oy];
                    const int node = (lx + ox) + M3_PITCH * (ly + oy);
#pragma unroll
                    for (int oz = 0; oz < 4; ++oz) {
                        const ValT w = wxy * a[2][oz];
                        const ValT* f = ring + ((cz - 1 + oz) & 3) * (M3_PL * LDA) + node;
#pragma unroll
                        for (int c = 0; c < LDA; ++c) acc[oz][c].add(w * f[c * M3_PL]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < LDA; ++c) {
                acc[0][c].merge(acc[1][c]);
                acc[2][c].merge(acc[3][c]);
                acc[0][c].merge(acc[2][c]);
            }
            ValT* u = up + (int64_t)p * lda + c0;
#pragma unroll
            for (int c = 0; c < LDA; ++c) u[c] = (ValT)acc[0][c].value();
        }
        __syncthreads();
        if (cz + 1 < z1) load_plane(cz + 3);
        __syncthreads();
    }
}

bool m2p3d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg) {
    if (m.dim != 3 || kernel != PPMG_KERNEL_MP4) return false;
    for (int d = 0; d < 3; ++d) {
        if (!m.periodic[d] || m.N[d] < 4) return false;
        if (fg.lo[d] != 0 || fg.ext[d] != m.N[d]) return false;
    }
    return true;
}

template <typename PosT, typename ValT>
static int m2p3d_typed(ppmg_plan* plan, const DevMesh& m, const PosT* xp, int lda, const ValT* field, ValT* up,
                       hipStream_t s) {
    dim3 grid((unsigned)((m.N[0] + M3_TT - 1) / M3_TT), (unsigned)((m.N[1] + M3_TT - 1) / M3_TT),
              (unsigned)((m.N[2] + M3_TZ - 1) / M3_TZ));
    for (int c0 = 0; c0 < lda; c0 += 4) {
        int L = lda - c0 < 4 ? lda - c0 : 4;
#define TL(LL) hipLaunchKernelGGL((k_m3_tile<LL, PosT, ValT>), grid, dim3(M3_THREADS), 0, s, m, plan->start.as<int>(), \
                                  plan->perm.as<int>(), xp, field, up, lda, c0)
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
int m2p3d_fast(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp, int lda, const void* field,
               void* up, hipStream_t s) {
    int64_t nbad = 0;
    if (bin_particles(plan, m, prec, np, xp, s, &nbad)) return PPMG_ERROR_FATAL;
    if (nbad > 0) return 1;
    mark(plan, s);
    if (prec == PPMG_PREC_F64)
        return m2p3d_typed<double, double>(plan, m, (const double*)xp, lda, (const double*)field, (double*)up, s);
    if (prec == PPMG_PREC_F32C)
        return m2p3d_typed<double, float>(plan, m, (const double*)xp, lda, (const float*)field, (float*)up, s);
    return m2p3d_typed<float, float>(plan, m, (const float*)xp, lda, (const float*)field, (float*)up, s);
}

}  // namespace ppmg
