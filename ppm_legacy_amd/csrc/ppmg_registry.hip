// ppmg_registry.hip -- PPM-compatible global state (topo_id / mesh_id) over the
// core C ABI, with host-array staging for callers that pass host memory (the
// Fortran modules).  Mirrors what ppm_init (ppm_init.f:161-434) and
// ppm_topo_mkfield (ppm_topo_mkfield.f:316-1040) leave behind on one process:
// one sub covering the domain, istart = 1, nnodes = Nm.
#include <cstring>
#include <map>
#include <mutex>

#include "ppmg_internal.h"
#include "ppmg_sort.h"

namespace ppmg {

struct Topo {
    ppmg_mesh mesh;
    int ghost[3];
};

static std::mutex g_mu;
static std::map<int, Topo> g_topos;
static int g_next_topo = 1;
static bool g_initialized = false;
static int g_prec_mode = PPMG_PREC_F64;

// Per-thread staging buffers and plan of the registry layer.
struct Staging {
    ppmg_plan* plan = nullptr;
    DevBuf xp, up, field, w[3], conv_up, conv_field, lp, lh;
    ~Staging() {
        DevBuf* b[] = {&xp, &up, &field, &w[0], &w[1], &w[2], &conv_up, &conv_field, &lp, &lh};
        for (DevBuf* x : b) x->release();
        if (plan) ppmg_plan_destroy(plan);
    }
};
static thread_local Staging g_stage;

static int current_info() { return ppmg_last_error(nullptr, nullptr, 0); }

static int reg_fail(int sev, int code, const char* msg) {
    set_error(sev, code, msg);
    return sev;
}

static int lookup(int topo_id, int mesh_id, Topo* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_initialized) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_NOINIT, "Please call ppm_init first!");
    auto it = g_topos.find(topo_id);
    if (it == g_topos.end()) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "topo_id is invalid!");
    if (mesh_id != 1) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "mesh_id is invalid!");
    *out = it->second;
    return 0;
}

static ppmg_plan* stage_plan() {
    if (!g_stage.plan) ppmg_plan_create(&g_stage.plan);
    return g_stage.plan;
}

// Host arrays are copied to device scratch; device arrays pass through.
static int stage_in(DevBuf& buf, const void* src, size_t bytes, int mem, const void** dev) {
    if (mem == PPMG_MEM_DEVICE || bytes == 0) {
        *dev = src;
        return 0;
    }
    if (buf.ensure(bytes)) return PPMG_ERROR_FATAL;
    PPMG_HIP(hipMemcpy(buf.p, src, bytes, hipMemcpyHostToDevice));
    *dev = buf.p;
    return 0;
}

static int stage_out_ptr(DevBuf& buf, void* dst, size_t bytes, int mem, void** dev) {
    if (mem == PPMG_MEM_DEVICE) {
        *dev = dst;
        return 0;
    }
    if (buf.ensure(bytes > 0 ? bytes : 1)) return PPMG_ERROR_FATAL;
    *dev = buf.p;
    return 0;
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_f2d(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

static int convert(const void* src, void* dst, int64_t n, bool to_float) {
    if (n <= 0) return 0;
    unsigned g = (unsigned)((n + 255) / 256);
    if (to_float) hipLaunchKernelGGL(k_d2f, dim3(g), dim3(256), 0, 0, (const double*)src, (float*)dst, n);
    else hipLaunchKernelGGL(k_f2d, dim3(g), dim3(256), 0, 0, (const float*)src, (double*)dst, n);
    PPMG_HIP(hipGetLastError());
    return 0;
}

static ppmg_field padded_field(const Topo& t, int lda, const int* ghost, void* data) {
    ppmg_field f;
    f.data = data;
    f.lda = lda;
    for (int d = 0; d < 3; ++d) {
        if (d < t.mesh.dim) {
            f.lo[d] = -ghost[d];
            f.ext[d] = t.mesh.nm[d] + 2 * ghost[d];
        } else {
            f.lo[d] = 0;
            f.ext[d] = 1;
        }
    }
    return f;
}

static int64_t field_count(const ppmg_field& f) { return (int64_t)f.lda * f.ext[0] * f.ext[1] * f.ext[2]; }

}  // namespace ppmg

using namespace ppmg;

extern "C" {

int ppmg_init(int dim, int prec, int tolexp, int debug) {
    (void)tolexp;
    clear_error();
    if (dim != 2 && dim != 3) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dim must be 2 or 3");
    if (prec != PPMG_PREC_F64 && prec != PPMG_PREC_F32)
        return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "prec must be ppm_kind_double or ppm_kind_single");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return reg_fail(PPMG_ERROR_FATAL, PPMG_ERR_DEVICE, "no HIP device visible");
    ppmg_set_debug(debug);
    std::lock_guard<std::mutex> lk(g_mu);
    g_initialized = true;
    return 0;
}

int ppmg_finalize(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_topos.clear();
    g_initialized = false;
    return 0;
}

int ppmg_set_precision_mode(int prec) {
    if (prec != PPMG_PREC_F64 && prec != PPMG_PREC_F32C)
        return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "precision mode must be F64 or F32C");
    g_prec_mode = prec;
    return 0;
}

int ppmg_mktopo_field(int dim, const double* min_phys, const double* max_phys, const int* bcdef, const int* nm,
                      const int* ghostsize, int* topo_id, int* mesh_id) {
    clear_error();
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_initialized) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_NOINIT, "Please call ppm_init first!");
    if (dim != 2 && dim != 3) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dim must be 2 or 3");
    Topo t;
    std::memset(&t, 0, sizeof(t));
    t.mesh.dim = dim;
    for (int d = 0; d < 3; ++d) {
        t.mesh.nm[d] = d < dim ? nm[d] : 2;
        t.mesh.min_phys[d] = d < dim ? min_phys[d] : 0.0;
        t.mesh.max_phys[d] = d < dim ? max_phys[d] : 1.0;
        t.ghost[d] = d < dim && ghostsize ? ghostsize[d] : 0;
    }
    for (int f = 0; f < 6; ++f) t.mesh.bcdef[f] = f < 2 * dim ? bcdef[f] : 1;
    for (int d = 0; d < dim; ++d)
        if (nm[d] < 2 || !(max_phys[d] > min_phys[d]))
            return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "invalid mesh or domain");
    int id = g_next_topo++;
    g_topos[id] = t;
    *topo_id = id;
    *mesh_id = 1;
    return 0;
}

int ppmg_topo_destroy(int topo_id) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_topos.erase(topo_id);
    return 0;
}

int ppmg_topo_info(int topo_id, int* dim, int* nm, double* min_phys, double* max_phys, int* bcdef) {
    clear_error();
    Topo t;
    if (lookup(topo_id, 1, &t)) return current_info();
    if (dim) *dim = t.mesh.dim;
    for (int d = 0; d < 3; ++d) {
        if (nm) nm[d] = t.mesh.nm[d];
        if (min_phys) min_phys[d] = t.mesh.min_phys[d];
        if (max_phys) max_phys[d] = t.mesh.max_phys[d];
    }
    if (bcdef)
        for (int f = 0; f < 6; ++f) bcdef[f] = t.mesh.bcdef[f];
    return 0;
}

int ppmg_impose_part_bc_reg(void* xp, int np, int topo_id, int prec, int mem) {
    clear_error();
    Topo t;
    if (lookup(topo_id, 1, &t)) return current_info();
    if (np <= 0) return 0;
    const bool single = prec == PPMG_PREC_F32;
    const size_t bytes = (single ? 4 : 8) * (size_t)np * t.mesh.dim;
    const void* dxp = nullptr;
    if (stage_in(g_stage.xp, xp, bytes, mem, &dxp)) return current_info();
    int rc = ppmg_impose_part_bc(&t.mesh, single ? PPMG_PREC_F32 : PPMG_PREC_F64, np, const_cast<void*>(dxp), nullptr);
    if (rc) return rc;
    if (mem == PPMG_MEM_HOST) PPMG_HIP(hipMemcpy(xp, dxp, bytes, hipMemcpyDeviceToHost));
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_rmsh_create_part_reg(const void* field_up, int lda, int topo_id, int mesh_id, const int* ghostsize,
                              const double* cutoff, const double* cutoff_weights, int resetpos, int maxnp,
                              int* np_out, void* xp, void* up, const void* field_wp, int lda2, void* wp,
                              int prec, int mem) {
    clear_error();
    Topo t;
    if (lookup(topo_id, mesh_id, &t)) return current_info();
    if (!ghostsize || !np_out) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize or np_out is NULL");
    int dim = t.mesh.dim;
    int g[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) g[d] = ghostsize[d];
    const bool single = prec == PPMG_PREC_F32;
    const size_t esz = single ? 4 : 8;
    ppmg_field f = padded_field(t, lda, g, nullptr);
    int64_t fcount = field_count(f);
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    const void* dfield = nullptr;
    void *dxp = nullptr, *dup = nullptr;
    if (stage_in(g_stage.field, field_up, esz * (size_t)fcount, mem, &dfield)) return current_info();
    if (stage_out_ptr(g_stage.xp, xp, esz * (size_t)(maxnp > 0 ? maxnp : 1) * dim, mem, &dxp)) return current_info();
    if (stage_out_ptr(g_stage.up, up, esz * (size_t)(maxnp > 0 ? maxnp : 1) * lda, mem, &dup)) return current_info();
    f.data = const_cast<void*>(dfield);
    ppmg_field fw{};
    void* dwp = nullptr;
    if (field_wp) {
        if (lda2 < 1) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda2 must be >= 1");
        fw = padded_field(t, lda2, g, nullptr);
        const void* dw = nullptr;
        if (stage_in(g_stage.conv_field, field_wp, esz * (size_t)field_count(fw), mem, &dw)) return current_info();
        if (stage_out_ptr(g_stage.w[0], wp, esz * (size_t)(maxnp > 0 ? maxnp : 1) * lda2, mem, &dwp))
            return current_info();
        fw.data = const_cast<void*>(dw);
    }
    int64_t n = 0;
    int rc = ppmg_rmsh_create_part(&t.mesh, single ? PPMG_PREC_F32 : PPMG_PREC_F64, &f, cutoff, cutoff_weights,
                                   resetpos, maxnp, &n, dxp, dup, field_wp ? &fw : nullptr, dwp, plan, nullptr);
    *np_out = (int)n;
    if (rc && !(maxnp == 0 && n > 0)) return rc;
    clear_error();
    if (mem == PPMG_MEM_HOST && maxnp > 0) {
        int64_t w = n < maxnp ? n : maxnp;
        if (w > 0) {
            PPMG_HIP(hipMemcpy(up, dup, esz * (size_t)w * lda, hipMemcpyDeviceToHost));
            if (field_wp) PPMG_HIP(hipMemcpy(wp, dwp, esz * (size_t)w * lda2, hipMemcpyDeviceToHost));
            if (resetpos) PPMG_HIP(hipMemcpy(xp, dxp, esz * (size_t)w * dim, hipMemcpyDeviceToHost));
        }
    }
    PPMG_HIP(hipDeviceSynchronize());
    return n > maxnp && maxnp > 0 ? reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ALLOC, "maxnp smaller than the new particles") : 0;
}

int ppmg_interp_p2m(const void* xp, int np, const void* up, int lda, int topo_id, int mesh_id, int kernel,
                    const int* ghostsize, void* field_up, int prec, int mem) {
    clear_error();
    // ppm_interp_p2m.f:338-367: with ppm_debug > 0 the initialisation and Np
    // checks come first, and Np == 0 returns before field_up is touched (no
    // zeroing, no staging, no copy back).
    if (debug_level() > 0) {
        {
            std::lock_guard<std::mutex> lk(g_mu);
            if (!g_initialized) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_NOINIT, "Please call ppm_init first!");
        }
        if (np < 0) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "particles must be specified");
        if (np == 0) return 0;
        if (kernel < 1 || kernel > 4) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "wrong kernel definition");
    }
    Topo t;
    if (lookup(topo_id, mesh_id, &t)) return current_info();
    if (!ghostsize) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize is NULL");
    if (np < 0) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "particles must be specified");
    if (lda < 1) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda must be >= 1");
    int dim = t.mesh.dim;
    int g[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) g[d] = ghostsize[d];
    const bool single = prec == PPMG_PREC_F32;
    const int form = single ? PPMG_PREC_F32 : g_prec_mode;
    const size_t esz = single ? 4 : 8;       // caller's element size
    ppmg_field f = padded_field(t, lda, g, field_up);
    int64_t fcount = field_count(f);
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    const void *dxp = nullptr, *dup = nullptr;
    void* dfield = nullptr;
    if (stage_in(g_stage.xp, xp, esz * (size_t)np * dim, mem, &dxp)) return current_info();
    if (stage_in(g_stage.up, up, esz * (size_t)np * lda, mem, &dup)) return current_info();
    if (stage_out_ptr(g_stage.field, field_up, esz * (size_t)fcount, mem, &dfield)) return current_info();
    if (form == PPMG_PREC_F32C) {
        // double caller arrays, float arithmetic: convert on the device
        if (g_stage.conv_up.ensure(sizeof(float) * (size_t)(np > 0 ? np : 1) * lda) ||
            g_stage.conv_field.ensure(sizeof(float) * (size_t)fcount))
            return current_info();
        if (convert(dup, g_stage.conv_up.p, (int64_t)np * lda, true)) return current_info();
        f.data = g_stage.conv_field.p;
        int rc = ppmg_p2m(&t.mesh, kernel, form, np, dxp, lda, g_stage.conv_up.p, &f, plan, nullptr);
        if (rc) return rc;
        if (convert(g_stage.conv_field.p, dfield, fcount, false)) return current_info();
    } else {
        f.data = dfield;
        int rc = ppmg_p2m(&t.mesh, kernel, form, np, dxp, lda, dup, &f, plan, nullptr);
        if (rc) return rc;
    }
    if (mem == PPMG_MEM_HOST) PPMG_HIP(hipMemcpy(field_up, dfield, esz * (size_t)fcount, hipMemcpyDeviceToHost));
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_interp_m2p(const void* xp, int np, void* up, int lda, int topo_id, int mesh_id, int kernel,
                    const int* ghostsize, const void* field_up, int prec, int mem) {
    clear_error();
    Topo t;
    if (lookup(topo_id, mesh_id, &t)) return current_info();
    if (!ghostsize) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize is NULL");
    if (np < 0) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "particles must be specified");
    if (lda < 1) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda must be >= 1");
    int dim = t.mesh.dim;
    int g[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) g[d] = ghostsize[d];
    const bool single = prec == PPMG_PREC_F32;
    const int form = single ? PPMG_PREC_F32 : g_prec_mode;
    const size_t esz = single ? 4 : 8;
    ppmg_field f = padded_field(t, lda, g, nullptr);
    int64_t fcount = field_count(f);
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    const void *dxp = nullptr, *dfield = nullptr;
    void* dup = nullptr;
    if (stage_in(g_stage.xp, xp, esz * (size_t)np * dim, mem, &dxp)) return current_info();
    if (stage_in(g_stage.field, field_up, esz * (size_t)fcount, mem, &dfield)) return current_info();
    if (stage_out_ptr(g_stage.up, up, esz * (size_t)np * lda, mem, &dup)) return current_info();
    if (form == PPMG_PREC_F32C) {
        if (g_stage.conv_up.ensure(sizeof(float) * (size_t)(np > 0 ? np : 1) * lda) ||
            g_stage.conv_field.ensure(sizeof(float) * (size_t)fcount))
            return current_info();
        if (convert(dfield, g_stage.conv_field.p, fcount, true)) return current_info();
        f.data = g_stage.conv_field.p;
        int rc = ppmg_m2p(&t.mesh, kernel, form, np, dxp, lda, &f, g_stage.conv_up.p, plan, nullptr);
        if (rc) return rc;
        if (convert(g_stage.conv_up.p, dup, (int64_t)np * lda, false)) return current_info();
    } else {
        f.data = const_cast<void*>(dfield);
        int rc = ppmg_m2p(&t.mesh, kernel, form, np, dxp, lda, &f, dup, plan, nullptr);
        if (rc) return rc;
    }
    if (mem == PPMG_MEM_HOST && np > 0) PPMG_HIP(hipMemcpy(up, dup, esz * (size_t)np * lda, hipMemcpyDeviceToHost));
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_rmsh_comp_weights_reg(const void* xp, int np, int topo_id, int mesh_id, int kernel, void* wx1, void* wx2,
                               void* wx3, int prec, int mem) {
    clear_error();
    Topo t;
    if (lookup(topo_id, mesh_id, &t)) return current_info();
    int dim = t.mesh.dim;
    const bool single = prec == PPMG_PREC_F32;
    const size_t esz = single ? 4 : 8;
    const int form = single ? PPMG_PREC_F32 : PPMG_PREC_F64;
    const int ks = kernel == PPMG_KERNEL_MP4 ? 4 : 2;
    const void* dxp = nullptr;
    void* dw[3] = {nullptr, nullptr, nullptr};
    void* hw[3] = {wx1, wx2, wx3};
    if (stage_in(g_stage.xp, xp, esz * (size_t)np * dim, mem, &dxp)) return current_info();
    for (int d = 0; d < dim; ++d)
        if (stage_out_ptr(g_stage.w[d], hw[d], esz * (size_t)np * ks, mem, &dw[d])) return current_info();
    int rc = ppmg_rmsh_comp_weights(&t.mesh, kernel, form, np, dxp, dw[0], dw[1], dim == 3 ? dw[2] : nullptr, nullptr);
    if (rc) return rc;
    if (mem == PPMG_MEM_HOST && np > 0)
        for (int d = 0; d < dim; ++d) PPMG_HIP(hipMemcpy(hw[d], dw[d], esz * (size_t)np * ks, hipMemcpyDeviceToHost));
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_rmsh_remesh_reg(const void* xp, int np, const void* up, int lda, int topo_id, int mesh_id, int kernel,
                         const int* ghostsize, void* field_up, const void* wx1, const void* wx2, const void* wx3,
                         int prec, int mem) {
    clear_error();
    Topo t;
    if (lookup(topo_id, mesh_id, &t)) return current_info();
    if (!ghostsize) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize is NULL");
    if (np < 0 || lda < 1) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "Np must be >= 0 and lda >= 1");
    int dim = t.mesh.dim;
    const int ks = kernel == PPMG_KERNEL_MP4 ? 2 : 1;
    int g[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        g[d] = ghostsize[d];
        if (g[d] < ks - 1)   // ppm_rmsh_remesh.f:340-347
            return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize too small for this kernel");
    }
    const bool single = prec == PPMG_PREC_F32;
    const size_t esz = single ? 4 : 8;
    const int form = single ? PPMG_PREC_F32 : PPMG_PREC_F64;
    ppmg_field f = padded_field(t, lda, g, nullptr);
    int64_t fcount = field_count(f);
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    const void *dxp = nullptr, *dup = nullptr, *dw[3] = {nullptr, nullptr, nullptr};
    const void* hw[3] = {wx1, wx2, wx3};
    void* dfield = nullptr;
    if (stage_in(g_stage.xp, xp, esz * (size_t)np * dim, mem, &dxp)) return current_info();
    if (stage_in(g_stage.up, up, esz * (size_t)np * lda, mem, &dup)) return current_info();
    for (int d = 0; d < dim; ++d)
        if (stage_in(g_stage.w[d], hw[d], esz * (size_t)np * 2 * ks, mem, &dw[d])) return current_info();
    if (stage_out_ptr(g_stage.field, field_up, esz * (size_t)fcount, mem, &dfield)) return current_info();
    f.data = dfield;
    int rc = ppmg_rmsh_remesh(&t.mesh, kernel, form, np, dxp, lda, dup, dw[0], dw[1], dim == 3 ? dw[2] : nullptr, &f,
                              plan, nullptr);
    if (rc) return rc;
    if (mem == PPMG_MEM_HOST) PPMG_HIP(hipMemcpy(field_up, dfield, esz * (size_t)fcount, hipMemcpyDeviceToHost));
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_map_field_ghost_reg(void* field_up, int lda, int topo_id, int mesh_id, const int* ghostsize, int maptype,
                             int prec, int mem) {
    // a copy / add of stored values: the double arrays take the FP64 form whatever the precision mode
    return ppmg_map_field_ghost_kind_reg(field_up, lda, topo_id, mesh_id, ghostsize, maptype,
                                         prec == PPMG_PREC_F32 ? PPMG_KIND_S : PPMG_KIND_D, nullptr, mem);
}

int ppmg_map_field_ghost_kind_reg(void* field_up, int lda, int topo_id, int mesh_id, const int* ghostsize,
                                  int maptype, int kind, const int* mask, int mem) {
    clear_error();
    Topo t;
    if (lookup(topo_id, mesh_id, &t)) return current_info();
    if (!ghostsize || !field_up) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize or field_up is NULL");
    if (lda < 1) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "lda must be >0");
    size_t esz;
    switch (kind) {
    case PPMG_KIND_D: esz = 8; break;
    case PPMG_KIND_S: case PPMG_KIND_I: case PPMG_KIND_L: esz = 4; break;
    case PPMG_KIND_DC: esz = 16; break;
    case PPMG_KIND_SC: esz = 8; break;
    default: return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "unknown element kind");
    }
    int g[3] = {0, 0, 0};
    for (int d = 0; d < t.mesh.dim; ++d) {
        g[d] = ghostsize[d];
        if (g[d] < 0) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "ghostsize must be >=0 in all dimensions");
    }
    ppmg_field f = padded_field(t, lda, g, nullptr);
    const size_t bytes = esz * (size_t)field_count(f);
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    void* dfield = nullptr;
    if (stage_out_ptr(g_stage.field, field_up, bytes, mem, &dfield)) return current_info();
    if (mem == PPMG_MEM_HOST) PPMG_HIP(hipMemcpy(dfield, field_up, bytes, hipMemcpyHostToDevice));
    const void* dmask = nullptr;
    if (mask && stage_in(g_stage.lh, mask, sizeof(int) * (size_t)(field_count(f) / lda), mem, &dmask))
        return current_info();
    f.data = dfield;
    int rc = ppmg_map_field_ghost_kind(&t.mesh, kind, &f, g, maptype, (const int*)dmask, plan, nullptr);
    if (rc) return rc;
    if (mem == PPMG_MEM_HOST) PPMG_HIP(hipMemcpy(field_up, dfield, bytes, hipMemcpyDeviceToHost));
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_util_sort_reg(int dim, void* xp, int np, const void* xmin, const void* xmax, const int* ncells, int* npbx,
                       int prec, int mem) {
    clear_error();
    if (!ncells || (np > 0 && !xp)) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "xp or Nm is NULL");
    if (dim != 2 && dim != 3) return reg_fail(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "dim must be 2 or 3");
    const bool single = prec == PPMG_PREC_F32;
    const int form = single ? PPMG_PREC_F32 : PPMG_PREC_F64;
    const size_t esz = single ? 4 : 8;
    int64_t nbox = 1;
    for (int d = 0; d < dim; ++d) nbox *= ncells[d] > 0 ? ncells[d] : 1;
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    void* dxp = nullptr;
    int* dnp = nullptr;
    if (stage_out_ptr(g_stage.xp, xp, esz * (size_t)(np > 0 ? np : 0) * dim, mem, &dxp)) return current_info();
    if (mem == PPMG_MEM_HOST && np > 0) PPMG_HIP(hipMemcpy(dxp, xp, esz * (size_t)np * dim, hipMemcpyHostToDevice));
    if (npbx) {
        void* p = nullptr;
        if (stage_out_ptr(g_stage.lh, npbx, sizeof(int) * (size_t)nbox, mem, &p)) return current_info();
        dnp = (int*)p;
    }
    int rc = ppmg_util_sort(dim, form, np, dxp, xmin, xmax, ncells, dnp, plan, nullptr);
    if (rc) return rc;
    if (mem == PPMG_MEM_HOST) {
        if (np > 0) PPMG_HIP(hipMemcpy(xp, dxp, esz * (size_t)np * dim, hipMemcpyDeviceToHost));
        if (npbx) PPMG_HIP(hipMemcpy(npbx, dnp, sizeof(int) * (size_t)nbox, hipMemcpyDeviceToHost));
    }
    PPMG_HIP(hipDeviceSynchronize());
    return 0;
}

int ppmg_util_rank_reg(int dim, const void* xp, int np, const void* xmin, const void* xmax, const int* ncells,
                       const int* ngl, int* lpdx, int* lhbx, int* icount, int prec, int mem) {
    clear_error();
    const bool single = prec == PPMG_PREC_F32;
    const size_t esz = single ? 4 : 8;
    int64_t nbox = 1;
    for (int d = 0; d < dim; ++d) nbox *= (int64_t)ncells[d] + ngl[d] + ngl[dim + d];
    ppmg_plan* plan = stage_plan();
    if (!plan) return current_info();
    const void* dxp = nullptr;
    void *dlp = nullptr, *dlh = nullptr;
    if (stage_in(g_stage.xp, xp, esz * (size_t)np * dim, mem, &dxp)) return current_info();
    if (stage_out_ptr(g_stage.lp, lpdx, sizeof(int) * (size_t)(np > 0 ? np : 1), mem, &dlp)) return current_info();
    if (stage_out_ptr(g_stage.lh, lhbx, sizeof(int) * (size_t)(nbox + 1), mem, &dlh)) return current_info();
    int64_t ic = 0;
    int rc = ppmg_util_rank(dim, single ? PPMG_PREC_F32 : PPMG_PREC_F64, np, dxp, xmin, xmax, ncells, ngl, (int*)dlp,
                            (int*)dlh, &ic, plan, nullptr);
    if (rc) return rc;
    if (mem == PPMG_MEM_HOST) {
        if (ic > 0) PPMG_HIP(hipMemcpy(lpdx, dlp, sizeof(int) * (size_t)ic, hipMemcpyDeviceToHost));
        PPMG_HIP(hipMemcpy(lhbx, dlh, sizeof(int) * (size_t)(nbox + 1), hipMemcpyDeviceToHost));
    }
    PPMG_HIP(hipDeviceSynchronize());
    *icount = (int)ic;
    return 0;
}

}  // extern "C"
