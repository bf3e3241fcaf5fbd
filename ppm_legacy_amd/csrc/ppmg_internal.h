// ppmg_internal.h -- shared device/host definitions of libppmg (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstring>
#include <string>

#include "ppmg.h"

namespace ppmg {

// Mesh as the kernels see it.  One sub covering the domain; node i (0-based)
// at minv + i*h; N = cells per axis = nm - 1; nbin = bin cells per axis: N on
// periodic axes (INT(x0) == N is the periodic image of cell 0), N + 1 on
// non-periodic ones (x == max_phys is inside the closed upper face).
struct DevMesh {
    int dim;
    int N[3];
    int nbin[3];
    int periodic[3];     // both faces periodic (bcdef 1)
    int upper_open[3];   // reference's sub test: x < max required on this face
    double minv[3];
    double maxv[3];
    double dxi[3];       // inverse spacing of the operation (see spacing_*)
};

struct FieldGeom {
    int lda;
    int lo[3];
    int ext[3];
    int swap[3];         // periodic axis storing both node 0 and node N: ghost_get node map
};

// Kahan-compensated float accumulator; value() = s - c.
struct KahanF {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float v) {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s - c; }
    // fold another compensated partial sum into this one
    __device__ __forceinline__ void merge(const KahanF& o) {
        add(o.s);
        add(-o.c);
    }
};

struct PlainD {
    double s;
    __device__ __forceinline__ void init() { s = 0.0; }
    __device__ __forceinline__ void add(double v) { s += v; }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ void merge(const PlainD& o) { s += o.s; }
};

template <typename ValT> struct AccOf;
template <> struct AccOf<double> { using type = PlainD; };
template <> struct AccOf<float> { using type = KahanF; };

// Cell index and fractional offset as the reference computes them: x0 is
// rounded first, x0 = (x - min)*dxi, then INT truncation and frac = x0 - INT(x0)
// (ppm_interp_p2m.f:800-822).  Contraction is off so that no FMA fuses the
// product into the subtraction (that would skip the rounding of x0 and move
// frac by up to one ulp of x0, visible in single precision).
template <typename T>
__device__ __forceinline__ T cell_offset(T x, T mn, T dxi, int& i) {
#pragma clang fp contract(off)
    T x0 = (x - mn) * dxi;
    i = (int)x0;
    return x0 - (T)i;
}

template <typename T>
__device__ __forceinline__ T scaled_coord(T x, T mn, T dxi) {
#pragma clang fp contract(off)
    return (x - mn) * dxi;
}

// ------------------------------------------------------ z-chunk choice --
// Workgroups of `kernel` (launched with `threads` threads) resident on the
// whole device at once, from the occupancy API (cached by the caller).
template <typename K>
inline int resident_workgroups(K kernel, int threads) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) return 1;
    return cus * per_cu > 0 ? cus * per_cu : 1;
}

// Cell planes per workgroup for the z-marching tile kernels.  A workgroup
// spends ~(tz + halo) plane steps; the launch takes ceil(workgroups / slots)
// such rounds.  Picks the chunk with the fewest plane steps overall (larger
// chunk on ties), so small meshes still fill the device and large ones march
// long chunks with little halo overhead.
inline int pick_z_chunk(int64_t tiles_xy, int nz, int slots, int halo) {
    int best = 32;
    int64_t best_cost = -1;
    for (int tz = 8; tz <= 128; tz *= 2) {
        int64_t nwg = tiles_xy * (int64_t)((nz + tz - 1) / tz);
        int64_t rounds = (nwg + slots - 1) / slots;
        int64_t cost = rounds * (int64_t)((tz < nz ? tz : nz) + halo);
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best = tz;
        }
    }
    return best;
}

// Cell planes per workgroup for the gathers (M2P, adjoint), whose z halo costs
// only ring loads: the longest chunk (8..128) that still gives `per_slot`
// workgroups per resident slot, so that workgroups of crowded tiles are short
// and the device balances them dynamically.  At 512^3 the clustered M2P takes
// 9.9 ms with 16-plane chunks against 18.3 ms with the 128 that pick_z_chunk's
// model of lockstep rounds prefers; the lattice's is unchanged (2.9 ms).
inline int pick_z_chunk_balanced(int64_t tiles_xy, int nz, int slots, int per_slot) {
    int best = 8;
    for (int tz = 8; tz <= 128; tz *= 2)
        if (tiles_xy * (int64_t)((nz + tz - 1) / tz) >= (int64_t)per_slot * slots) best = tz;
    return best;
}

// ----------------------------------------------------- XCD tile order --
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2.  Tile kernels use a 1D grid and map
// workgroup b to tile (b % 8) * q + min(b % 8, r) + b / 8 (q, r: tiles per
// XCD and remainder), so each XCD walks a contiguous block of tiles in x-fastest
// order and the halos that neighbouring tiles share are served from one L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t ntiles) {
    const int64_t x = b % NUM_XCD, j = b / NUM_XCD;
    const int64_t q = ntiles / NUM_XCD, r = ntiles % NUM_XCD;
    return x * q + (x < r ? x : r) + j;
}

// ghost_get copies shifted images of the sub's real nodes only
// (ppm_map_field_ghost_init.f:234-447, ppm_mesh_block_intersect.f:207-279):
// on a mesh with non-periodic axes it fills a node only when the node's
// coordinate on every non-periodic axis is a real node 0..N.  Such a node takes
// the periodic node map on the periodic axes; any other node keeps what is
// stored there, on every axis.  True when node v of axis d allows the map.
__device__ __forceinline__ bool real_on_open_axis(const DevMesh& m, int d, int v) {
    return m.periodic[d] || (v >= 0 && v <= m.N[d]);
}

// ------------------------------------------------------ DPP row shifts --
// Move a value K lanes up (row_shr: lane l reads lane l-K) or down (row_shl:
// lane l reads lane l+K) inside each 16-lane row; bound_ctrl makes lanes whose
// source lies outside the row read 0.  Doubles move as two dwords.
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
template <int K>
__device__ __forceinline__ double row_shl(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x100 + K, 0xF, 0xF, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x100 + K, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <int K>
__device__ __forceinline__ float row_shl(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + K, 0xF, 0xF, true));
}

// Value of lane ex + D of the same 16-lane row (D > 0: row_shl, D < 0: row_shr).
template <int D, typename T>
__device__ __forceinline__ T lane_from(T v) {
    if constexpr (D > 0) return row_shl<D>(v);
    else if constexpr (D < 0) return row_shr<-D>(v);
    else return v;
}

// Stored node read for node v along axis d, as the general path reads it:
// on a periodic axis the compact layout wraps, and a stored box holding node
// 0 and node N reads with ghost_get's node map (ppm_map_field_ghost_get.f:
// 257-280: node 0 and the ghosts below it take node v+N, node N and the ghosts
// above it v-N); an open axis reads node v itself (beside a periodic box the
// MIX form reads some nodes through raw_node instead).  Stencils reach -1..N+2;
// the ring of a tile also loads columns beyond that which no stencil reads:
// they wrap (periodic) or are clamped into the box (open).
template <bool OPEN>
__device__ __forceinline__ int stored_node(const DevMesh& m, const FieldGeom& fg, int d, int v) {
    const int N = m.N[d];
    if (OPEN && !m.periodic[d]) return min(max(v, fg.lo[d]), fg.lo[d] + fg.ext[d] - 1);
    if (fg.swap[d] && v > -N && v < 2 * N) return v <= 0 ? v + N : (v >= N ? v - N : v);
    v %= N;
    return v < 0 ? v + N : v;
}

// Orders one wavefront's LDS writes before its later LDS reads by other lanes
// of the same wavefront (a wavefront's LDS instructions execute in order; this
// keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------- weights --
// M'4 on the nodes INT(x0)-1 .. INT(x0)+2 (0-based) in the 3D unrolled form
// of ppm_interp_p2m.f:824-853 (same polynomials as ppm_interp_m2p.f).
template <typename T>
__device__ __forceinline__ void mp4_unrolled(T frac, T a[4]) {
    T x10 = frac + T(1);
    T x11 = x10 - T(1);
    T x12 = x10 - T(2);
    T x13 = x10 - T(3);
    a[0] = T(2) + (T(-4) + (T(2.5) - T(0.5) * x10) * x10) * x10;
    a[1] = T(1) + (T(-2.5) + T(1.5) * x11) * (x11 * x11);
    a[2] = T(1) + (T(-2.5) - T(1.5) * x12) * (x12 * x12);
    a[3] = T(2) + (T(4) + (T(2.5) + T(0.5) * x13) * x13) * x13;
}

// The same four weights in factored form, 11 operations instead of 16: with
// t = frac and s = 1 - t, W(1+t) = -t s^2/2, W(t) = 1 + t^2 (1.5 t - 2.5),
// W(s) = 1 + s^2 (1.5 s - 2.5) and W(2-t) = -t^2 s/2.  Equal to mp4_unrolled
// up to rounding (a few ulps); the tiled kernels use it.
template <typename T>
__device__ __forceinline__ void mp4_factored(T t, T a[4]) {
    const T s = T(1) - t, t2 = t * t, s2 = s * s;
    a[0] = (T(-0.5) * t) * s2;
    a[1] = T(1) + t2 * (T(1.5) * t - T(2.5));
    a[2] = T(1) + s2 * (T(1.5) * s - T(2.5));
    a[3] = (T(-0.5) * s) * t2;
}

// BSP2 weights in the reference's form (ppm_interp_p2m.f:3077-3104):
// 1 - t and 1 + (t - 1); bsp2_at is weight o alone.
template <typename T>
__device__ __forceinline__ void bsp2_weights(T t, T a[2]) {
    a[0] = T(1) - t;
    a[1] = T(1) + (t - T(1));
}
template <typename T>
__device__ __forceinline__ T bsp2_at(T t, int o) {
    return o == 0 ? T(1) - t : T(1) + (t - T(1));
}

// d/dt of mp4_factored's weights, factored the same way (equal to
// mp4_unrolled_deriv up to rounding): s (1.5 t - 0.5), t (4.5 t - 5),
// s (5 - 4.5 s), t (1.5 t - 1).
template <typename T>
__device__ __forceinline__ void mp4_factored_deriv(T t, T d[4]) {
    const T s = T(1) - t;
    d[0] = s * (T(1.5) * t - T(0.5));
    d[1] = t * (T(4.5) * t - T(5));
    d[2] = s * (T(5) - T(4.5) * s);
    d[3] = t * (T(1.5) * t - T(1));
}

// Weight o (0..3) of mp4_unrolled alone, without branches.  With
// t = (frac + 1) - o the four unrolled polynomials are the inner form
// 1 + (-2.5 + 1.5|t|) t^2 (o = 1, 2) and the outer form
// 2 + (-4 + (2.5 - 0.5|t|)|t|)|t| (o = 0, 3), up to exact sign changes.
template <typename T>
__device__ __forceinline__ T mp4_at(T frac, int o) {
    const T s = fabs((frac + T(1)) - T(o));
    const T in = T(1) + (T(-2.5) + T(1.5) * s) * (s * s);
    const T out = T(2) + (T(-4) + (T(2.5) - T(0.5) * s) * s) * s;
    return (o == 1 || o == 2) ? in : out;
}

// Stencil geometry of the tiled kernels per axis: KS nodes per particle (4:
// M'4, 2: BSP2); a particle of cell c reaches nodes c - B .. c - B + KS - 1.
template <int KS> struct Stencil {
    static constexpr int H = KS - 1;       // halo cells a tile reads on the low side of an axis
    static constexpr int B = KS == 4 ? 1 : 0;
};

// The KS 1D weights of a particle at fraction t of its cell, and weight o alone.
template <int KS, typename T>
__device__ __forceinline__ void tile_weights(T t, T (&a)[KS]) {
    if constexpr (KS == 4) mp4_factored<T>(t, a);
    else bsp2_weights<T>(t, a);
}
template <int KS, typename T>
__device__ __forceinline__ T tile_weight_at(T t, int o) {
    if constexpr (KS == 4) return mp4_at<T>(t, o);
    else return bsp2_at<T>(t, o);
}

// M'4 by distance branch, the 2D form of ppm_interp_p2m.f:3620-3651.
template <typename T>
__device__ __forceinline__ T mp4_branch(T r) {
    if (r < T(1)) return T(1) - (r * r) * (T(2.5) - T(1.5) * r);
    return T(2) + (T(-4) + (T(2.5) - T(0.5) * r) * r) * r;
}

// d/ds of the four M'4 weights (s = x0 - node) in the same unrolled variables.
template <typename T>
__device__ __forceinline__ void mp4_unrolled_deriv(T frac, T d[4]) {
    T x10 = frac + T(1);
    T x11 = x10 - T(1);
    T x12 = x10 - T(2);
    T x13 = x10 - T(3);
    d[0] = T(-4) + (T(5) - T(1.5) * x10) * x10;
    d[1] = (T(-5) + T(4.5) * x11) * x11;
    d[2] = (T(-5) - T(4.5) * x12) * x12;
    d[3] = T(4) + (T(5) + T(1.5) * x13) * x13;
}

// 1D weights of one axis.  KS = 4 (M'4) or 2 (BSP2); DIM selects the
// reference's formula family (3D unrolled vs 2D branch form for M'4).
template <int KS, int DIM, typename T>
__device__ __forceinline__ void weights_1d(T frac, T a[KS]) {
    if constexpr (KS == 4) {
        if constexpr (DIM == 3) {
            mp4_unrolled<T>(frac, a);
        } else {
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                T r = frac - T(o - 1);
                a[o] = mp4_branch<T>(r < T(0) ? -r : r);
            }
        }
    } else {
        // ppm_interp_p2m.f:3077-3104: a10 = 1 - x10, a11 = 1 + (x10 - 1)
        a[0] = T(1) - frac;
        a[1] = T(1) + (frac - T(1));
    }
}

// -------------------------------------------------------------- host side --
struct Status {
    int info = 0;
    int code = 0;
    std::string msg;
};

void set_error(int severity, int code, const std::string& msg);
void clear_error();
int debug_level();

#define PPMG_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::ppmg::set_error(PPMG_ERROR_FATAL, PPMG_ERR_DEVICE,                         \
                              std::string(#call) + ": " + hipGetErrorString(e_));        \
            return PPMG_ERROR_FATAL;                                                     \
        }                                                                                \
    } while (0)

// Growable device buffer owned by a plan.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;   // carved from a bound Arena: never freed
    int ensure(size_t n);
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Fixed device memory of an allocation-free call.  While an arena is bound on
// the calling thread, DevBuf::ensure carves buffers from it in bump order
// (256-byte aligned) instead of calling hipMalloc, and a request past its end
// fails with PPMG_ERR_ALLOC.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
};
void bind_arena(Arena* a);   // nullptr unbinds
bool arena_bound();          // an arena is bound on the calling thread

}  // namespace ppmg

namespace ppmg {
// Two descriptors that bin and weigh particles identically (bitwise equal
// origin and inverse spacing).
inline bool same_mesh(const DevMesh& a, const DevMesh& b) {
    if (a.dim != b.dim) return false;
    for (int d = 0; d < 3; ++d) {
        if (a.N[d] != b.N[d] || a.nbin[d] != b.nbin[d] || a.periodic[d] != b.periodic[d] ||
            a.upper_open[d] != b.upper_open[d])
            return false;
        if (std::memcmp(&a.minv[d], &b.minv[d], sizeof(double)) != 0 ||
            std::memcmp(&a.maxv[d], &b.maxv[d], sizeof(double)) != 0 ||
            std::memcmp(&a.dxi[d], &b.dxi[d], sizeof(double)) != 0)
            return false;
    }
    return true;
}
}  // namespace ppmg

// Opaque plan: scratch of the binning and the forward record for the adjoint.
struct ppmg_plan {
    ppmg::DevBuf key, slot, count, start, perm, wsorted, usorted, scan_tmp, err, misc, misc2, gaps, multi,
        multi_n, ghost, table, rhist;
    ppmg::DevBuf rkey[2], rval[2];   // ping-pong pairs of the radix-sort binning of crowded input
    int64_t np = 0;
    int64_t ncells = 0;
    bool sorted_input = false;   // last binning found the input already in cell order
    int64_t crowded_cells = 0;   // estimate (1-in-64 sample) of the last binning's cells holding nine or more particles
    // The last binning left a slot table (ppmg::TABLE_SLOTS int32 planes of
    // ncells: particle r of cell c at table[r * ncells + c], with the byte
    // counts in `count`) instead of `start` and `perm`; materialize_heads()
    // builds those two from it.
    bool bin_table = false;
    int table_skip = 0;          // binnings to run without the table after one overflowed it
    // The last binning found a lattice in cell order (particle p in cell p,
    // np == ncells) on lattice_mesh in precision lattice_prec; bin_lattice:
    // the current binning is that lattice, taken without a pass and not yet
    // checked (only a reader that checks every particle's cell accepts it).
    bool lattice_pred = false;
    int lattice_prec = 0;
    ppmg::DevMesh lattice_mesh{};
    bool bin_lattice = false;
    bool lattice_checked = false;   // ... and a gather has checked it
    int lattice_skip = 0;           // binnings to run without the lattice after a check failed
    bool perm_identity = false;  // the permutation of the last binning is the identity ...
    bool perm_written = true;    // ... and has (or has not yet) been written into `perm`
    hipStream_t bin_stream = nullptr;   // stream of the last binning
    const void* bin_xp = nullptr;       // positions, precision and mesh of the last binning
    int bin_prec = 0;
    ppmg::DevMesh bin_mesh{};
    bool keys_written = true;    // false: the last binning (cell-ordered input) wrote no keys
    // ppmg_m2p_prebin: the binning of a coming ppmg_m2p (its positions,
    // count, precision and mesh, and the particles it found outside the sub)
    bool prebinned = false;
    const void* prebin_xp = nullptr;
    int64_t prebin_np = 0, prebin_nbad = 0;
    int prebin_prec = 0;
    ppmg::DevMesh prebin_mesh{};
    int dim = 0, kernel = 0, prec = 0;
    ppmg::DevMesh mesh{};
    bool have_forward = false;
    hipEvent_t ev[8] = {};
    bool ev_init = false;
    bool timing = false;
    int nev = 0;                 // events recorded by the last call
    hipStream_t ev_stream = nullptr;
    double last_ms[8] = {};
};

namespace ppmg {
// Phase boundary marker of a timed plan (no-op unless timing is on).
inline void mark(ppmg_plan* plan, hipStream_t s, bool first = false) {
    if (!plan || !plan->timing) return;
    if (!plan->ev_init) {
        for (auto& e : plan->ev) (void)hipEventCreate(&e);
        plan->ev_init = true;
    }
    if (first) plan->nev = 0;
    if (plan->nev < 8) (void)hipEventRecord(plan->ev[plan->nev++], s);
    plan->ev_stream = s;
}
}  // namespace ppmg
