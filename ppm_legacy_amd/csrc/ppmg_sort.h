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

int exclusive_scan(const int* in, int64_t n, int* out, DevBuf& tmp, int64_t total, hipStream_t s);
int fill_zero(void* p, size_t bytes, hipStream_t s);
// Bins np particles into the plan (key, perm, start); *nbad = particles outside the sub.
int bin_particles(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp,
                  hipStream_t s, int64_t* nbad);
// The binning leaves the identity permutation implicit when the input is
// already in cell order.  Tiled kernels take tile_perm(): nullptr then, and
// they read particle f at sorted position f; other readers of plan->perm call
// materialize_perm() first.
inline const int* tile_perm(const ppmg_plan* plan) { return plan->perm_identity ? nullptr : plan->perm.as<int>(); }
int materialize_perm(ppmg_plan* plan, hipStream_t s);
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
