// ppmg_api.hip -- C ABI of libppmg: argument checks, mesh descriptors, plans,
// the PPM-compatible topo_id/mesh_id registry and host staging.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "ppmg_internal.h"
#include "ppmg_interp.h"
#include "ppmg_sort.h"

namespace ppmg {

static thread_local Status g_status;
static std::atomic<int> g_debug{0};

void set_error(int severity, int code, const std::string& msg) {
    g_status.info = severity;
    g_status.code = code;
    g_status.msg = msg;
    if (g_debug.load() > 0) std::fprintf(stderr, "ppmg: %04d %s\n", code, msg.c_str());
}
void clear_error() {
    g_status.info = 0;
    g_status.code = 0;
    g_status.msg.clear();
}
int debug_level() { return g_debug.load(); }

// Fault injection for the error path of device allocations: with
// PPMG_FAIL_ALLOC_ABOVE=<bytes> set, every new scratch allocation larger than
// that many bytes fails as a refused hipMalloc would (read only when a buffer
// has to grow, so calls that reuse their scratch never look at it).
static bool injected_alloc_failure(size_t n) {
    const char* e = std::getenv("PPMG_FAIL_ALLOC_ABOVE");
    if (!e || !*e) return false;
    return n > (size_t)std::strtoull(e, nullptr, 10);
}

static thread_local Arena* g_arena = nullptr;
void bind_arena(Arena* a) { g_arena = a; }
bool arena_bound() { return g_arena != nullptr; }

int DevBuf::ensure(size_t n) {
    if (n <= bytes && p) return 0;
    release();
    size_t want = n < 256 ? 256 : n;
    if (Arena* a = g_arena) {
        want = (want + 255) & ~(size_t)255;
        if (want > a->cap - a->used) {
            set_error(PPMG_ERROR_FATAL, PPMG_ERR_ALLOC,
                      "allocation-free call: its scratch needs more than the fixed " + std::to_string(a->cap) +
                          " bytes (" + std::to_string(a->used) + " used, " + std::to_string(want) +
                          " more requested)");
            return PPMG_ERROR_FATAL;
        }
        p = a->base + a->used;
        a->used += want;
        bytes = want;
        borrowed = true;
        return 0;
    }
    hipError_t e = injected_alloc_failure(want) ? hipErrorOutOfMemory : hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        set_error(PPMG_ERROR_FATAL, PPMG_ERR_ALLOC, std::string("hipMalloc of device scratch failed: ") + hipGetErrorString(e));
        return PPMG_ERROR_FATAL;
    }
    bytes = want;
    return 0;
}
void DevBuf::release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
}

static int fail(int severity, int code, const std::string& msg) {
    set_error(severity, code, msg);
    return severity;
}

// Spacing as P2M and remesh compute it: REAL(MK) throughout
// (ppm_interp_p2m.f:479-489, ppm_rmsh_remesh.f:600-606).
static double spacing_mk(const ppmg_mesh* m, int d, int prec) {
    if (prec == PPMG_PREC_F32) {
        float len = (float)m->max_phys[d] - (float)m->min_phys[d];
        float dx = len / (float)(m->nm[d] - 1);
        return (double)(1.0f / dx);
    }
    double len = m->max_phys[d] - m->min_phys[d];
    double dx = len / (double)(m->nm[d] - 1);
    return 1.0 / dx;
}

// Spacing as M2P and comp_weights compute it: len_phys, dx, dxi declared
// default REAL (single) even in the double specifics
// (ppm_interp_m2p.f:221-222,481-487; ppm_rmsh_comp_weights.f:146-147).
static double spacing_single(const ppmg_mesh* m, int d, int prec) {
    if (prec == PPMG_PREC_F32) {
        float len = (float)m->max_phys[d] - (float)m->min_phys[d];
        float dx = len / (float)(m->nm[d] - 1);
        return (double)(1.0f / dx);
    }
    float len = (float)(m->max_phys[d] - m->min_phys[d]);
    float dx = (float)((double)len / (double)(m->nm[d] - 1));
    float dxi = (float)(1.0 / (double)dx);
    return (double)dxi;
}

enum class Spacing { MK, Single };

static int make_mesh(const ppmg_mesh* in, int prec, Spacing sp, DevMesh* m) {
    if (!in) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "mesh descriptor is NULL");
    if (in->dim != 2 && in->dim != 3) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dim must be 2 or 3");
    if (prec != PPMG_PREC_F64 && prec != PPMG_PREC_F32C && prec != PPMG_PREC_F32)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "unknown precision form");
    std::memset(m, 0, sizeof(*m));
    m->dim = in->dim;
    for (int d = 0; d < 3; ++d) {
        m->N[d] = 1;
        m->nbin[d] = 1;
        m->dxi[d] = 1.0;
        m->maxv[d] = 1.0;
    }
    for (int d = 0; d < in->dim; ++d) {
        if (in->nm[d] < 2) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Nm must be >= 2 on every axis");
        if (!(in->max_phys[d] > in->min_phys[d]))
            return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "max_phys must exceed min_phys");
        bool plo = in->bcdef[2 * d] == 1, phi = in->bcdef[2 * d + 1] == 1;
        if (plo != phi)
            return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "a periodic axis needs periodic bcdef on both faces");
        m->N[d] = in->nm[d] - 1;
        m->periodic[d] = plo ? 1 : 0;
        m->upper_open[d] = phi ? 1 : 0;
        m->nbin[d] = plo ? m->N[d] : m->N[d] + 1;
        m->minv[d] = in->min_phys[d];
        m->maxv[d] = in->max_phys[d];
        m->dxi[d] = sp == Spacing::MK ? spacing_mk(in, d, prec) : spacing_single(in, d, prec);
    }
    int64_t ncells = (int64_t)m->nbin[0] * m->nbin[1] * m->nbin[2];
    if (ncells >= (int64_t)1 << 31) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31 cells");
    return 0;
}

static int make_field(const ppmg_field* in, const DevMesh& m, FieldGeom* f) {
    if (!in || !in->data) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field descriptor or data is NULL");
    if (in->lda < 1) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda must be >= 1");
    f->lda = in->lda;
    for (int d = 0; d < 3; ++d) {
        if (d < m.dim) {
            if (in->ext[d] < 1) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field extent must be >= 1");
            f->lo[d] = in->lo[d];
            f->ext[d] = in->ext[d];
            f->swap[d] = m.periodic[d] && in->lo[d] <= 0 && in->lo[d] + in->ext[d] - 1 >= m.N[d];
        } else {
            f->lo[d] = 0;
            f->ext[d] = 1;
            f->swap[d] = 0;
        }
    }
    return 0;
}

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }


static int check_plan(ppmg_plan* plan) {
    if (!plan) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "plan is NULL");
    return 0;
}

}  // namespace ppmg

using namespace ppmg;

extern "C" {

int ppmg_last_error(int* code, char* msg, int msglen) {
    if (code) *code = g_status.code;
    if (msg && msglen > 0) {
        std::strncpy(msg, g_status.msg.c_str(), (size_t)msglen - 1);
        msg[msglen - 1] = 0;
    }
    return g_status.info;
}

int ppmg_set_debug(int debug) {
    g_debug.store(debug);
    return 0;
}

int ppmg_device_sync(void* stream) {
    PPMG_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int ppmg_plan_create(ppmg_plan** plan) {
    clear_error();
    if (!plan) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "plan pointer is NULL");
    *plan = new (std::nothrow) ppmg_plan();
    if (!*plan) return fail(PPMG_ERROR_FATAL, PPMG_ERR_ALLOC, "plan allocation failed");
    const char* env = std::getenv("PPMG_TIMING");
    (*plan)->timing = env && env[0] && env[0] != '0';
    return 0;
}

int ppmg_plan_destroy(ppmg_plan* plan) {
    if (!plan) return 0;
    DevBuf* bufs[] = {&plan->key, &plan->slot, &plan->count, &plan->start, &plan->perm, &plan->wsorted,
                      &plan->usorted, &plan->scan_tmp, &plan->err, &plan->misc, &plan->misc2, &plan->gaps, &plan->multi, &plan->multi_n, &plan->ghost,
                      &plan->table, &plan->rhist, &plan->rkey[0], &plan->rkey[1], &plan->rval[0], &plan->rval[1]};
    for (DevBuf* b : bufs) b->release();
    if (plan->ev_init)
        for (auto& e : plan->ev) (void)hipEventDestroy(e);
    delete plan;
    return 0;
}

int ppmg_plan_set_timing(ppmg_plan* plan, int on) {
    if (!plan) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "plan is NULL");
    plan->timing = on != 0;
    return 0;
}

int ppmg_plan_timings(ppmg_plan* plan, double* ms, int n) {
    if (!plan || !ms) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "plan or ms is NULL");
    if (!plan->timing || plan->nev < 2) return 0;
    PPMG_HIP(hipEventSynchronize(plan->ev[plan->nev - 1]));
    int k = 0;
    for (int i = 0; i + 1 < plan->nev && k < n; ++i, ++k) {
        float t = 0.f;
        PPMG_HIP(hipEventElapsedTime(&t, plan->ev[i], plan->ev[i + 1]));
        ms[k] = t;
        plan->last_ms[k] = t;
    }
    return k;
}

int ppmg_map_field_ghost(const ppmg_mesh* mesh, int prec, const ppmg_field* field, const int* ghostsize, int maptype,
                         ppmg_plan* plan, void* stream) {
    return ppmg_map_field_ghost_kind(mesh, prec == PPMG_PREC_F64 ? PPMG_KIND_D : PPMG_KIND_S, field, ghostsize,
                                     maptype, nullptr, plan, stream);
}

int ppmg_map_field_ghost_kind(const ppmg_mesh* mesh, int kind, const ppmg_field* field, const int* ghostsize,
                              int maptype, const int* mask, ppmg_plan* plan, void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (maptype != PPMG_MAP_GHOST_GET && maptype != PPMG_MAP_GHOST_PUT)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "maptype must be ghost_get or ghost_put");
    if (!ghostsize) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize is NULL");
    if (kind != PPMG_KIND_D && kind != PPMG_KIND_S && kind != PPMG_KIND_I && kind != PPMG_KIND_L &&
        kind != PPMG_KIND_DC && kind != PPMG_KIND_SC)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "unknown element kind");
    // the mesh geometry is read for its node counts and boundary conditions only
    const int prec = kind == PPMG_KIND_S || kind == PPMG_KIND_SC ? PPMG_PREC_F32 : PPMG_PREC_F64;
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::MK, &m)) return g_status.info;
    FieldGeom fg;
    if (make_field(field, m, &fg)) return g_status.info;
    int g[3] = {0, 0, 0};
    for (int d = 0; d < m.dim; ++d) {
        g[d] = ghostsize[d];
        if (g[d] < 0) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize must be >=0 in all dimensions");
        // the stored box must hold nodes 1-g .. Nm+g (0-based -g .. N+g)
        if (fg.lo[d] > -g[d] || fg.lo[d] + fg.ext[d] - 1 < m.N[d] + g[d])
            return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field does not hold the mesh and its ghost layers");
    }
    if (map_field_ghost(m, kind, fg, field->data, g, maptype == PPMG_MAP_GHOST_GET, mask, plan->ghost,
                        as_stream(stream)))
        return g_status.info;
    return 0;
}

int ppmg_plan_sort_view(const ppmg_plan* plan, const int** key, const int** perm, const int** start, int64_t* np,
                        int64_t* ncells) {
    if (!plan) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "plan is NULL");
    if ((perm || start) && plan->bin_table) {
        // the binning left a slot table: build the heads and the permutation
        ppmg_plan* w = const_cast<ppmg_plan*>(plan);
        if (materialize_heads(w, w->bin_stream)) return g_status.info;
        PPMG_HIP(hipStreamSynchronize(w->bin_stream));
    }
    if (perm && plan->perm_identity && !plan->perm_written) {
        // the binning left the identity permutation implicit: write it now
        ppmg_plan* w = const_cast<ppmg_plan*>(plan);
        if (materialize_perm(w, w->bin_stream)) return g_status.info;
        PPMG_HIP(hipStreamSynchronize(w->bin_stream));
    }
    if (key && !plan->keys_written) {
        // input binned in cell order left the keys unwritten: compute them now
        ppmg_plan* w = const_cast<ppmg_plan*>(plan);
        if (materialize_keys(w, w->bin_stream)) return g_status.info;
        PPMG_HIP(hipStreamSynchronize(w->bin_stream));
    }
    if (key) *key = plan->key.as<int>();
    if (perm) *perm = plan->perm.as<int>();
    if (start) *start = plan->start.as<int>();
    if (np) *np = plan->np;
    if (ncells) *ncells = plan->ncells;
    return 0;
}

int ppmg_p2m(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, int lda, const void* up,
             const ppmg_field* field, ppmg_plan* plan, void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Only Mp4 and Bsp2 are available.");
    if (np < 0) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "particles must be specified");
    if (lda < 1) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda must be >= 1");
    if (np > 0 && (!xp || !up)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "xp or up is NULL");
    if (np >= ((int64_t)1 << 31)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31-1 particles");
    // ppm_interp_p2m.f:359-367: with ppm_debug > 0 and Np == 0 nothing is touched
    if (np == 0 && debug_level() > 0) return 0;
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::MK, &m)) return g_status.info;
    FieldGeom fg;
    if (make_field(field, m, &fg)) return g_status.info;
    if (fg.lda != lda) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field lda differs from lda");
    hipStream_t s = as_stream(stream);
    int64_t nbad = 0;
    plan->have_forward = false;
    // PPMG_GENERAL=1 forces the general node-gather path (used to cross-check the fast path)
    const char* gen = std::getenv("PPMG_GENERAL");
    const bool force_general = gen && gen[0] && gen[0] != '0';
    mark(plan, s, true);
    if (bin_particles(plan, m, prec, np, xp, s, &nbad,
                      !force_general &&
                          (p2m3d_table_applicable(m, kernel, prec, fg) || p2m2d_table_applicable(m, kernel, fg))))
        return g_status.info;
    if (nbad > 0)
        return fail(PPMG_ERROR_FATAL, PPMG_ERR_PART_UNASS,
                    "MAJOR PROBLEM: " + std::to_string(nbad) + " particles lie in no sub");
    mark(plan, s);
    if (!force_general && p2m3d_fast_applicable(m, kernel, prec, fg)) {
        if (p2m3d_fast(plan, m, kernel, prec, np, xp, fg, up, field->data, nullptr, s)) return g_status.info;
    } else if (!force_general && p2m2d_fast_applicable(m, kernel, fg)) {
        if (p2m2d_fast(plan, m, kernel, prec, xp, fg, up, field->data, s)) return g_status.info;
    } else if (p2m_general(plan, m, kernel, prec, np, xp, lda, up, fg, field->data, nullptr, s)) {
        return g_status.info;
    }
    plan->mesh = m;
    plan->dim = m.dim;
    plan->kernel = kernel;
    plan->prec = prec;
    plan->np = np;
    plan->have_forward = true;
    return 0;
}

int ppmg_m2p(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, int lda,
             const ppmg_field* field, void* up, ppmg_plan* plan, void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Only Mp4 and Bsp2 are available.");
    if (np < 0) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "particles must be specified");
    if (lda < 1) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda must be >= 1");
    if (np > 0 && (!xp || !up)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "xp or up is NULL");
    if (np >= ((int64_t)1 << 31)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31-1 particles");
    if (np == 0) return 0;
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::Single, &m)) return g_status.info;
    FieldGeom fg;
    if (make_field(field, m, &fg)) return g_status.info;
    if (fg.lda != lda) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field lda differs from lda");
    hipStream_t s = as_stream(stream);
    // a binning ppmg_m2p_prebin left for exactly these particles is used once
    const bool prebinned = plan->prebinned && plan->prebin_xp == xp && plan->prebin_np == np &&
                           plan->prebin_prec == prec && same_mesh(plan->prebin_mesh, m);
    plan->prebinned = false;
    if (plan->err.ensure(sizeof(int) * 4) || fill_zero(plan->err.p, sizeof(int) * 4, s)) return g_status.info;
    mark(plan, s, true);
    const char* gen = std::getenv("PPMG_GENERAL");
    const bool force_general = gen && gen[0] && gen[0] != '0';
    const bool tiled3 = !force_general && m2p3d_fast_applicable(m, kernel, fg);
    const bool tiled2 = !force_general && !tiled3 && m2p2d_fast_applicable(m, kernel, fg);
    if (tiled3 || tiled2) {
        int rc = tiled3 ? m2p3d_fast(plan, m, kernel, prec, np, xp, fg, field->data, up, prebinned, s)
                        : m2p2d_fast(plan, m, kernel, prec, np, xp, fg, field->data, up, s);
        if (rc < 0) return g_status.info;
        if (rc == 0) {
            mark(plan, s);
            plan->have_forward = false;   // the binning now describes these particles
            return 0;
        }
    }
    // the binning of the tiled path counts particles outside the sub in err[0]: start from zero
    if (fill_zero(plan->err.p, sizeof(int) * 4, s)) return g_status.info;
    if (m2p_general(m, kernel, prec, np, xp, lda, fg, field->data, up, plan->err.as<int>(), s)) return g_status.info;
    mark(plan, s);
    int bad = 0;
    PPMG_HIP(hipMemcpyAsync(&bad, plan->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PPMG_HIP(hipStreamSynchronize(s));
    if (bad)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_FIELD_BOUNDS,
                    std::to_string(bad) + " particles read nodes outside the stored field");
    return 0;
}

int ppmg_m2p_prebin(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, ppmg_plan* plan,
                    void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    plan->prebinned = false;
    if (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Only Mp4 and Bsp2 are available.");
    if (np < 0) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "particles must be specified");
    if (np > 0 && !xp) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "xp is NULL");
    if (np >= ((int64_t)1 << 31)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31-1 particles");
    if (np == 0) return 0;
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::Single, &m)) return g_status.info;
    // only the 3D tiled gather (M'4 and BSP2 bin alike) reads a binning
    if (m.dim != 3) return 0;
    int64_t nbad = 0;
    plan->have_forward = false;
    // the tiled gather checks a lattice taken without a pass (m2p3d_fast)
    if (bin_particles(plan, m, prec, np, xp, as_stream(stream), &nbad, m2p3d_table_applicable(m),
                      m2p3d_table_applicable(m)))
        return g_status.info;
    plan->prebinned = true;
    plan->prebin_xp = xp;
    plan->prebin_np = np;
    plan->prebin_nbad = nbad;
    plan->prebin_prec = prec;
    plan->prebin_mesh = m;
    return 0;
}

int ppmg_p2m_adjoint(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, int lda,
                     const void* up, const ppmg_field* g, void* grad_up, void* grad_xp, ppmg_plan* plan,
                     void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (!plan->have_forward || plan->np != np || plan->kernel != kernel || plan->prec != prec)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT,
                    "adjoint needs the forward ppmg_p2m of the same particles on this plan");
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::MK, &m)) return g_status.info;
    if (m.dim != plan->dim) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dimension differs from the forward");
    // The adjoint decodes the forward's cell keys and heads: the mesh must be
    // the one the forward binned on (cells, bins, boundary conditions, origin
    // and spacing), else keys would be read with the wrong stride or past the
    // heads array.  Checked before any launch.
    if (!same_mesh(m, plan->mesh))
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "mesh differs from the one the forward binned on");
    FieldGeom fg;
    if (make_field(g, m, &fg)) return g_status.info;
    if (fg.lda != lda) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "cotangent lda differs from lda");
    if (np == 0) return 0;
    mark(plan, as_stream(stream), true);
    // PPMG_GENERAL=1 forces the general per-particle path (used to cross-check the tiled one)
    const char* gen = std::getenv("PPMG_GENERAL");
    const bool force_general = gen && gen[0] && gen[0] != '0';
    if (!force_general && adjoint3d_fast_applicable(m, kernel, fg)) {
        if (adjoint3d_fast(plan, m, prec, np, xp, lda, up, fg, g->data, grad_up, grad_xp, as_stream(stream)))
            return g_status.info;
    } else if (adjoint_general(plan, m, kernel, prec, np, xp, lda, up, fg, g->data, grad_up, grad_xp,
                               as_stream(stream))) {
        return g_status.info;
    }
    mark(plan, as_stream(stream));
    return 0;
}

int ppmg_util_rank(int dim, int prec, int64_t np, const void* xp, const void* xmin, const void* xmax,
                   const int* ncells, const int* ngl, int* lpdx, int* lhbx, int64_t* icount, ppmg_plan* plan,
                   void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (dim != 2 && dim != 3) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dim must be 2 or 3");
    if (np < 0) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Np must be >= 0");
    if (np >= ((int64_t)1 << 31)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31-1 particles");
    RankGeom g;
    std::memset(&g, 0, sizeof(g));
    g.dim = dim;
    for (int d = 0; d < 3; ++d) { g.ntot[d] = 1; g.nm[d] = 1; g.xmax[d] = 1; g.rdx[d] = 1; }
    for (int d = 0; d < dim; ++d) {
        if (ncells[d] < 1) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Nm must be >= 1");
        if (ngl[d] < 0 || ngl[dim + d] < 0) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Ngl must be >= 0");
        g.nm[d] = ncells[d];
        g.ngl_lo[d] = ngl[d];
        g.ntot[d] = ncells[d] + ngl[d] + ngl[dim + d];
        if (prec == PPMG_PREC_F32) {
            float mn = ((const float*)xmin)[d], mx = ((const float*)xmax)[d];
            g.xmin[d] = mn;
            g.xmax[d] = mx;
            g.rdx[d] = (double)((float)ncells[d] / (mx - mn));   // ppm_util_rank3d.f:353-355
        } else {
            double mn = ((const double*)xmin)[d], mx = ((const double*)xmax)[d];
            g.xmin[d] = mn;
            g.xmax[d] = mx;
            g.rdx[d] = (double)ncells[d] / (mx - mn);
        }
    }
    if ((int64_t)g.ntot[0] * g.ntot[1] * g.ntot[2] >= ((int64_t)1 << 31))
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31 boxes");
    if (rank_particles(plan, g, prec, np, xp, lpdx, lhbx, icount, as_stream(stream))) return g_status.info;
    return 0;
}

int ppmg_util_sort(int dim, int prec, int64_t np, void* xp, const void* xmin, const void* xmax, const int* ncells,
                   int* npbx, ppmg_plan* plan, void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (dim != 2 && dim != 3) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dim must be 2 or 3");
    if (!ncells) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Nm is NULL");
    int64_t nbox = 1;
    for (int d = 0; d < dim; ++d) nbox *= ncells[d] > 0 ? ncells[d] : 1;
    const int ngl[6] = {0, 0, 0, 0, 0, 0};
    if (plan->misc.ensure(sizeof(int) * (size_t)(np > 0 ? np : 1)) ||
        plan->misc2.ensure(sizeof(int) * (size_t)(nbox + 1)))
        return g_status.info;
    int* lp = plan->misc.as<int>();
    int* lh = plan->misc2.as<int>();
    int64_t icount = 0;
    int rc = ppmg_util_rank(dim, prec, np, xp, xmin, xmax, ncells, ngl, lp, lh, &icount, plan, stream);
    if (rc) return rc;
    // ppm_util_sort3d.f:226-232: every particle must lie in a box
    if (icount < np) return fail(PPMG_ERROR_ERROR, PPMG_ERR_PART_LOST, "Not all particles have been ranked");
    if (sort_rows(dim, prec, np, xp, lp, lh, nbox, npbx, plan->usorted, as_stream(stream))) return g_status.info;
    return 0;
}

int ppmg_impose_part_bc(const ppmg_mesh* mesh, int prec, int64_t np, void* xp, void* stream) {
    clear_error();
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::MK, &m)) return g_status.info;
    if (impose_bc_general(m, prec, np, xp, as_stream(stream))) return g_status.info;
    return 0;
}

int ppmg_rmsh_comp_weights(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, void* wx1,
                           void* wx2, void* wx3, void* stream) {
    clear_error();
    if (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "wrong kernel definition");
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::Single, &m)) return g_status.info;
    if (np > 0 && (!xp || !wx1 || !wx2 || (m.dim == 3 && !wx3)))
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "cannot compute weights into empty arrays");
    if (comp_weights_general(m, kernel, prec, np, xp, wx1, wx2, wx3, as_stream(stream))) return g_status.info;
    return 0;
}

int ppmg_rmsh_remesh(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, int lda,
                     const void* up, const void* wx1, const void* wx2, const void* wx3, const ppmg_field* field,
                     ppmg_plan* plan, void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (kernel != PPMG_KERNEL_MP4 && kernel != PPMG_KERNEL_BSP2)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "wrong kernel definition");
    if (np < 0 || lda < 1) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Np must be >= 0 and lda >= 1");
    if (np >= ((int64_t)1 << 31)) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "more than 2^31-1 particles");
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::MK, &m)) return g_status.info;
    if (np > 0 && (!wx1 || !wx2 || (m.dim == 3 && !wx3)))
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "cannot remesh with empty weights");
    FieldGeom fg;
    if (make_field(field, m, &fg)) return g_status.info;
    if (fg.lda != lda) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field lda differs from lda");
    hipStream_t s = as_stream(stream);
    int64_t nbad = 0;
    plan->have_forward = false;
    if (bin_particles(plan, m, prec, np, xp, s, &nbad)) return g_status.info;
    if (nbad > 0)
        return fail(PPMG_ERROR_FATAL, PPMG_ERR_PART_UNASS,
                    "MAJOR PROBLEM: " + std::to_string(nbad) + " particles lie in no sub");
    const void* tabs[3] = {wx1, wx2, wx3};
    // 3D M'4 on the tiled deposit with the caller's weights (PPMG_GENERAL=1
    // forces the general path, for cross-checks)
    const char* gen = std::getenv("PPMG_GENERAL");
    const bool force_general = gen && gen[0] && gen[0] != '0';
    if (!force_general && kernel == PPMG_KERNEL_MP4 && p2m3d_fast_applicable(m, kernel, prec, fg)) {
        if (p2m3d_fast(plan, m, kernel, prec, np, xp, fg, up, field->data, tabs, s)) return g_status.info;
    } else if (p2m_general(plan, m, kernel, prec, np, xp, lda, up, fg, field->data, tabs, s)) {
        return g_status.info;
    }
    return 0;
}

int ppmg_rmsh_create_part(const ppmg_mesh* mesh, int prec, const ppmg_field* field, const double* cutoff,
                          const double* cutoff_weights, int resetpos, int64_t maxnp, int64_t* np_out, void* xp,
                          void* up, const ppmg_field* slave, void* wp, ppmg_plan* plan, void* stream) {
    clear_error();
    if (check_plan(plan)) return PPMG_ERROR_ERROR;
    if (!cutoff || !np_out) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "cutoff or np_out is NULL");
    DevMesh m;
    if (make_mesh(mesh, prec, Spacing::MK, &m)) return g_status.info;
    FieldGeom fg;
    if (make_field(field, m, &fg)) return g_status.info;
    if (cutoff_weights && fg.lda > 16)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "cutoff weights support lda <= 16");
    for (int d = 0; d < m.dim; ++d) {
        int need_hi = m.upper_open[d] ? m.N[d] - 1 : m.N[d];
        if (fg.lo[d] > 0 || fg.lo[d] + fg.ext[d] - 1 < need_hi)
            return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "field does not store the real nodes");
    }
    if (maxnp > 0 && (!up || (resetpos && !xp)))
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "output arrays are NULL");
    FieldGeom fw;
    if (slave) {
        if (make_field(slave, m, &fw)) return g_status.info;
        for (int d = 0; d < m.dim; ++d) {
            int need_hi = m.upper_open[d] ? m.N[d] - 1 : m.N[d];
            if (fw.lo[d] > 0 || fw.lo[d] + fw.ext[d] - 1 < need_hi)
                return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "slave field does not store the real nodes");
        }
        if (maxnp > 0 && !wp) return fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "wp is NULL");
    }
    if (create_part_general(plan, m, prec, fg, field->data, cutoff, cutoff_weights, resetpos, maxnp, np_out, xp, up,
                            slave ? &fw : nullptr, slave ? slave->data : nullptr, wp, as_stream(stream)))
        return g_status.info;
    if (*np_out > maxnp)
        return fail(PPMG_ERROR_ERROR, PPMG_ERR_ALLOC, "output capacity smaller than the number of new particles");
    return 0;
}

}  // extern "C"
