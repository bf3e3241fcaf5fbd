// ppmg_interp.h -- launchers of the general interpolation path (ppmg_interp.hip)
// and of the 3D M'4 fast path (ppmg_p2m3d.hip).
#pragma once
#include "ppmg_internal.h"

namespace ppmg {

int p2m_general(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, int lda,
                const void* up, const FieldGeom& fg, void* field, const void* const* wtab, hipStream_t s);
int m2p_general(const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, int lda, const FieldGeom& fg,
                const void* field, void* up, int* err, hipStream_t s);
int adjoint_general(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, int lda,
                    const void* up, const FieldGeom& fg, const void* g, void* gu, void* gx, hipStream_t s);
int comp_weights_general(const DevMesh& m, int kernel, int prec, int64_t np, const void* xp, void* w1, void* w2,
                         void* w3, hipStream_t s);
int impose_bc_general(const DevMesh& m, int prec, int64_t np, void* xp, hipStream_t s);
int create_part_general(ppmg_plan* plan, const DevMesh& m, int prec, const FieldGeom& fg, const void* field,
                        const double* cutoff, const double* cw, int resetpos, int64_t maxnp, int64_t* np_out,
                        void* xp, void* up, const FieldGeom* fw, const void* wfield, void* wp, hipStream_t s);

// 3D tiled deposit (ppmg_p2m3d.hip), M'4 or BSP2: per axis (periodic or open)
// the compact layout of a periodic axis or a stored box holding the real range 0..N.
bool p2m3d_fast_applicable(const DevMesh& m, int kernel, int prec, const FieldGeom& fg);
// ... and it reads a slot-table binning (bin_particles' `table`) there.
bool p2m3d_table_applicable(const DevMesh& m, int kernel, int prec, const FieldGeom& fg);
// wtab: NULL (weights from the positions), or, for M'4 only, the remesh weight
// tables (four per particle and axis, ppm_rmsh_comp_weights' layout).
int p2m3d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp,
               const FieldGeom& fg, const void* up, void* field, const void* const* wtab, hipStream_t s);

// 2D tiled deposit (ppmg_p2m2d.hip), M'4 or BSP2: per axis the compact layout
// of a periodic axis or a stored box holding the real range 0..N.
bool p2m2d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg);
// ... and it reads a slot-table binning (bin_particles' `table`) there.
bool p2m2d_table_applicable(const DevMesh& m, int kernel, const FieldGeom& fg);
int p2m2d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, const void* xp, const FieldGeom& fg,
               const void* up, void* field, hipStream_t s);

// 3D tiled gather (ppmg_m2p3d.hip), M'4 or BSP2, same layouts as the deposit:
// bins the particles into the plan first (unless prebinned: the plan already
// holds their binning); returns 1 without touching up when a particle lies
// outside the sub.
bool m2p3d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg);
// ... and it reads a slot-table binning (bin_particles' `table`) there.
bool m2p3d_table_applicable(const DevMesh& m);
int m2p3d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp,
               const FieldGeom& fg, const void* field, void* up, bool prebinned, hipStream_t s);

// 2D tiled gather (ppmg_m2p2d.hip), M'4 or BSP2: periodic axes compact or a
// stored box holding 0..N, open axes a box holding every node a stencil
// reaches (no periodic box beside an open axis).  Bins first; returns 1
// without touching up when a particle lies outside the sub.
bool m2p2d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg);
int m2p2d_fast(ppmg_plan* plan, const DevMesh& m, int kernel, int prec, int64_t np, const void* xp,
               const FieldGeom& fg, const void* field, void* up, hipStream_t s);

// Tiled 3D M'4 adjoint (ppmg_adj3d.hip), on the forward's binning; periodic
// axes, each in the compact layout or a stored box holding 0..N.
bool adjoint3d_fast_applicable(const DevMesh& m, int kernel, const FieldGeom& fg);
int adjoint3d_fast(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp, int lda, const void* up,
                   const FieldGeom& fg, const void* g, void* gu, void* gx, hipStream_t s);

// Ghost mapping of a padded field (ppmg_ghost.hip).  A block is the part of
// one periodic image of the real nodes that the stored box holds: src is its
// first real node (1-based), size its extent, off the image shift.
struct GhostBlock {
    int src[3];
    int size[3];
    int off[3];
};
int ghost_blocks(const DevMesh& m, const int* g, GhostBlock* out);
// kind: PPMG_KIND_*; mask: one logical per stored node (nullptr: all nodes)
int map_field_ghost(const DevMesh& m, int kind, const FieldGeom& fg, void* field, const int* g, bool get,
                    const int* mask, DevBuf& scratch, hipStream_t s);

}  // namespace ppmg
