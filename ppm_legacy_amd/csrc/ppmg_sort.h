// ppmg_sort.h -- counting sort / ranking entry points shared inside libppmg.
#pragma once
#include "ppmg_internal.h"

namespace ppmg {

// Geometry of ppm_util_rank3d / rank2d (ppm_util_rank3d.f:309-355).
struct RankGeom {
    int dim;
    int nm[3];        // cells per axis (the reference's Nm)
    int ngl_lo[3];
    int ntot[3];      // nm + ngl_lo + ngl_hi (1 on unused axes)
    double xmin[3], xmax[3];
    double rdx[3];    // Nm / (xmax - xmin) in the position precision
};

// Byte-wide cell counts of the binning.  In every block of 256 cells, cell c
// sits in word c mod 64, byte (c / 64) mod 4: the 64 consecutive cells a
// wavefront of lattice-ordered particles hits fall in 64 different words, so
// the atomic counts of neighbouring lanes do not contend for one word.  The
// array is padded to whole blocks and zeroed, so reads need no bounds.
__host__ __device__ inline int64_t count_word(int64_t c) { return ((c >> 8) << 6) + (c & 63); }
__host__ __device__ inline unsigned count_shift(int64_t c) { return (unsigned)((c >> 6) & 3) * 8u; }

// Slot table of a binning (plan->bin_table): particle r (0 <= r < count) of
// cell c, in increasing particle index, is table[r * ncells + c].  A cell
// with more than TABLE_SLOTS particles sends the binning to the heads and
// permutation instead.  Eight slots hold every cell of a lattice jittered by
// less than half a cell and displaced by a locally uniform shift: a cell then
// meets at most two particles per axis.
constexpr int TABLE_SLOTS = 8;

// A slot-table binning as the tiled kernels read it.
struct BinTable {
    const uint32_t* count;
    const int* table;
    int64_t ncells;
};

// Cell key of one particle with the index arithmetic of ppm_interp_p2m.f:
// 800-822 (x0 = (x - min)*dxi, INT truncation) and the sub test of :535-561;
// -1 when the particle lies in no cell of the sub.
template <int DIM, typename PosT>
__device__ __forceinline__ int cell_key(const PosT (&x)[DIM], const DevMesh& m) {
    bool ok = true;
    int cell[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        PosT v = x[d];
        PosT mn = (PosT)m.minv[d];
        PosT mx = (PosT)m.maxv[d];
        if (!(v >= mn && v <= mx)) ok = false;
        if (!(v < mx || !m.upper_open[d])) ok = false;
        PosT x0 = scaled_coord<PosT>(v, mn, (PosT)m.dxi[d]);
        int i = ok ? (int)x0 : 0;
        if (m.periodic[d] && i >= m.N[d]) i -= m.N[d];
        if (i < 0 || i >= m.nbin[d]) ok = false;
        cell[d] = i;
    }
    return ok ? cell[0] + m.nbin[0] * (cell[1] + m.nbin[1] * cell[2]) : -1;
}

int exclusive_scan(const int* in, int64_t n, int* out, DevBuf& tmp, int64_t total, hipStream_t s);
int fill_zero(void* p, size_t bytes, hipStream_t s);
// Bins np particles into the plan (key, perm, start); *nbad = particles outside the sub.
// table: the caller reads a slot table (plan->bin_table) where the binning
// can leave one instead of the heads and the permutation.  lattice: the
// caller checks every particle's cell (err[6]), so a plan that predicts a
// lattice in cell order may take it without a pass (plan->bin_lattice).
int bin_particles(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp,
                  hipStream_t s, int64_t* nbad, bool table = false, bool lattice = false);
// Heads and permutation of a binning that left a slot table (no-op otherwise).
int materialize_heads(ppmg_plan* plan, hipStream_t s);
// The binning leaves the identity permutation implicit when the input is
// already in cell order.  Tiled kernels take tile_perm(): nullptr then, and
// they read particle f at sorted position f; other readers of plan->perm call
// materialize_perm() first.
inline const int* tile_perm(const ppmg_plan* plan) { return plan->perm_identity ? nullptr : plan->perm.as<int>(); }
int materialize_perm(ppmg_plan* plan, hipStream_t s);   // calls materialize_heads() first
// Input binned in cell order leaves the keys unwritten too; readers of
// plan->key call materialize_keys() first (it reads the positions of that
// binning, which must still hold them).
int materialize_keys(ppmg_plan* plan, hipStream_t s);
int rank_particles(ppmg_plan* plan, const RankGeom& g, int prec, int64_t np, const void* xp,
                   int* lpdx, int* lhbx, int64_t* icount, hipStream_t s);
// xp (np rows of dim) reordered by the 1-based ranking lpdx1; npbx (may be
// NULL) = particles per box from the 1-based heads lhbx1.
int sort_rows(int dim, int prec, int64_t np, void* xp, const int* lpdx1, const int* lhbx1, int64_t nbox, int* npbx,
              DevBuf& tmp, hipStream_t s);

}  // namespace ppmg
