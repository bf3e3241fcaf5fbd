// ppmg_noalloc.hip -- ppmg_p2m for callers that cannot allocate device memory.
//
// The binning scratch of ppmg_p2m (keys, slots, permutation, cell counts and
// heads, scan partials, error words) normally lives in a plan whose buffers
// grow by hipMalloc.  ppmg_p2m_noalloc runs the same code on the same kernels
// with every such buffer carved from one fixed device array, `scratch`, of
// PPMG_NOALLOC_SCRATCH_BYTES, so the call performs no device allocation or
// free.  Each GPU holds its own copy of the array until the process ends; its
// address on a device is looked up once, with that device current, and only
// that address is handed to kernels and copies.  Calls on one GPU share the
// array (the last call's contents stay there; the binning clears what must
// start at zero), so calls that overlap on one GPU race.
#include <cstring>
#include <mutex>
#include <string>

#include "ppmg_internal.h"

namespace ppmg {
namespace {

constexpr size_t NOALLOC_BYTES = (size_t)PPMG_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device).
int scratch_address(void** out) {
    int dev = 0;
    PPMG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error(PPMG_ERROR_FATAL, PPMG_ERR_DEVICE, "device ordinal beyond the allocation-free scratch table");
        return PPMG_ERROR_FATAL;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        PPMG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}

// Unbinds the arena however the call returns.
struct ArenaScope {
    explicit ArenaScope(Arena* a) { bind_arena(a); }
    ~ArenaScope() { bind_arena(nullptr); }
};

}  // namespace
}  // namespace ppmg

using namespace ppmg;

extern "C" int ppmg_p2m_noalloc(const ppmg_mesh* mesh, int kernel, int prec, int64_t np, const void* xp, int lda,
                                const void* up, const ppmg_field* field, void* stream) {
    clear_error();
    // Before any launch: the binning alone holds three int32 per particle
    // (key, arrival slot, permutation) and two per cell (counts, heads).
    if (mesh && (mesh->dim == 2 || mesh->dim == 3) && np > 0) {
        int64_t ncells = 1;
        for (int d = 0; d < mesh->dim; ++d) {
            const int64_t nm = mesh->nm[d] > 1 ? mesh->nm[d] : 2;
            ncells *= mesh->bcdef[2 * d] == 1 ? nm - 1 : nm;
        }
        const double need = 12.0 * (double)np + 8.0 * (double)ncells;
        if (need > (double)NOALLOC_BYTES) {
            set_error(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT,
                      std::to_string(np) + " particles on " + std::to_string(ncells) +
                          " cells need more scratch than the " + std::to_string(NOALLOC_BYTES) +
                          " bytes of the allocation-free P2M");
            return PPMG_ERROR_ERROR;
        }
    }
    void* base = nullptr;
    if (scratch_address(&base)) return PPMG_ERROR_FATAL;
    Arena arena;
    arena.base = static_cast<char*>(base);
    arena.cap = NOALLOC_BYTES;
    ppmg_plan plan;   // timing off: no events are created
    ArenaScope bound(&arena);
    // Every buffer the plan grows comes from the arena; the plan is dropped
    // without release (nothing in it was allocated).
    return ppmg_p2m(mesh, kernel, prec, np, xp, lda, up, field, &plan, stream);
}
