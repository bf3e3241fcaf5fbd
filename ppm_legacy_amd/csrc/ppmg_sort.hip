// ppmg_sort.hip -- stable counting sort of particles into cells (gfx950).
//
// Replaces the cell ranking of ppm_util_rank3d / rank2d
// (ppm_util_rank3d.f:309-532): count particles per cell, exclusive prefix
// over the cells, scatter in particle order.  On the GPU the count and the
// scatter use integer atomics, which place the particles of one cell in an
// arbitrary order; a per-cell pass then restores increasing particle index,
// so the permutation equals the reference's lpdx exactly.
#include <climits>
#include <cstdlib>

#include "ppmg_internal.h"
#include "ppmg_sort.h"

namespace ppmg {

// Binning status words: err[0] particles outside the sub, err[1] a descent
// seen, err[2] queued runs of empty cells, err[3] a byte count overflowed,
// err[4] a cell with more particles than the slot table holds, err[5] a
// particle p not in cell p, err[6] (the gather's check of a lattice
// binning) the same, err[8..71] sampled crowded cells.
constexpr int ERR_WORDS = 72;

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t& excl) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wid = t >> 6;
    // wavefront inclusive scan (64 lanes)
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
            int64_t s = sh[w];
            sh[w] = run;
            run += s;
        }
        sh[SCAN_BLOCK / 64] = run;
    }
    __syncthreads();
    excl = x - v + sh[wid];
    int64_t total = sh[SCAN_BLOCK / 64];
    __syncthreads();
    return total;
}

// Each thread owns SCAN_ITEMS consecutive elements of its tile, read (and in
// k_scan_down written) as 128-bit vectors when the tile is complete.
template <bool VEC>
__device__ __forceinline__ void load_items(const int* __restrict__ in, int64_t n, int64_t base, int (&v)[SCAN_ITEMS]) {
    if (VEC && base + SCAN_ITEMS <= n) {
        const int4* q = reinterpret_cast<const int4*>(in + base);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS / 4; ++k) {
            int4 t = q[k];
            v[4 * k] = t.x; v[4 * k + 1] = t.y; v[4 * k + 2] = t.z; v[4 * k + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) v[i] = base + i < n ? in[base + i] : 0;
    }
}

template <bool VEC>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const int* __restrict__ in, int64_t n,
                                                            int64_t* __restrict__ block_sums) {
    __shared__ int64_t sh[SCAN_BLOCK / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    load_items<VEC>(in, n, base, v);
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) s += v[i];
    int64_t excl;
    int64_t tot = block_exclusive_scan(s, sh, excl);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Exclusive scan of the block sums by one workgroup of SUMS_THREADS threads:
// each thread totals a contiguous segment, the segment totals are scanned
// across the workgroup, and each thread rewrites its segment as prefixes.
constexpr int SUMS_THREADS = 1024;
__global__ void __launch_bounds__(SUMS_THREADS) k_scan_sums(int64_t* __restrict__ sums, int64_t nb) {
    __shared__ int64_t sh[SUMS_THREADS / 64 + 1];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int64_t seg = (nb + SUMS_THREADS - 1) / SUMS_THREADS;
    const int64_t b = t * seg, e = b + seg < nb ? b + seg : nb;
    int64_t tot = 0;
    for (int64_t i0 = b; i0 < e; i0 += 8) {   // eight loads in flight at a time
        int64_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = i0 + j < e ? sums[i0 + j] : 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) tot += v[j];
    }
    // workgroup exclusive scan of the segment totals
    int64_t x = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int w = 0; w < SUMS_THREADS / 64; ++w) {
            int64_t v = sh[w];
            sh[w] = run;
            run += v;
        }
    }
    __syncthreads();
    int64_t run = x - tot + sh[wid];
    for (int64_t i0 = b; i0 < e; i0 += 8) {
        int64_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = i0 + j < e ? sums[i0 + j] : 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (i0 + j < e) sums[i0 + j] = run;
            run += v[j];
        }
    }
}

template <bool VEC>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_down(const int* __restrict__ in, int64_t n,
                                                          const int64_t* __restrict__ offs,
                                                          int* __restrict__ out, int64_t total_out) {
    __shared__ int64_t sh[SCAN_BLOCK / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int vals[SCAN_ITEMS];
    load_items<VEC>(in, n, base, vals);
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) s += vals[i];
    int64_t excl;
    block_exclusive_scan(s, sh, excl);
    int64_t run = offs[blockIdx.x] + excl;
    int o[SCAN_ITEMS];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        o[i] = (int)run;
        run += vals[i];
    }
    if (VEC && base + SCAN_ITEMS <= n) {
        int4* q = reinterpret_cast<int4*>(out + base);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS / 4; ++k) q[k] = make_int4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            if (base + i < n) out[base + i] = o[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BLOCK - 1) out[n] = (int)total_out;
}

int exclusive_scan(const int* in, int64_t n, int* out, DevBuf& tmp, int64_t total, hipStream_t s) {
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) nb = 1;
    if (tmp.ensure(sizeof(int64_t) * (size_t)nb)) return PPMG_ERROR_FATAL;
    int64_t* sums = tmp.as<int64_t>();
    // 128-bit loads and stores need 16-byte aligned arrays (caller arrays may not be)
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    if (vec) hipLaunchKernelGGL(k_scan_reduce<true>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums);
    else hipLaunchKernelGGL(k_scan_reduce<false>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SUMS_THREADS), 0, s, sums, nb);
    if (vec) hipLaunchKernelGGL(k_scan_down<true>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums, out, total);
    else hipLaunchKernelGGL(k_scan_down<false>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums, out, total);
    PPMG_HIP(hipGetLastError());
    return 0;
}

// Counts of the 16 cells base .. base+15 (base a multiple of 16): 16
// consecutive words of one block, one byte of each.
__device__ __forceinline__ void load_counts16(const uint32_t* __restrict__ words, int64_t base, int (&c)[16]) {
    const uint4* q = reinterpret_cast<const uint4*>(words + count_word(base));
    const unsigned sh = count_shift(base);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 v = q[k];
        c[4 * k] = (int)((v.x >> sh) & 255u);
        c[4 * k + 1] = (int)((v.y >> sh) & 255u);
        c[4 * k + 2] = (int)((v.z >> sh) & 255u);
        c[4 * k + 3] = (int)((v.w >> sh) & 255u);
    }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce_u8(const uint32_t* __restrict__ in,
                                                               int64_t* __restrict__ block_sums) {
    __shared__ int64_t sh[SCAN_BLOCK / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int c[SCAN_ITEMS];
    load_counts16(in, base, c);
    int sum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) sum += c[i];
    int64_t excl;
    const int64_t tot = block_exclusive_scan(sum, sh, excl);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_down_u8(const uint32_t* __restrict__ in, int64_t n,
                                                             const int64_t* __restrict__ offs, int* __restrict__ out,
                                                             int64_t total_out) {
    __shared__ int64_t sh[SCAN_BLOCK / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int c[SCAN_ITEMS];
    load_counts16(in, base, c);
    int sum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) sum += c[i];
    int64_t excl;
    block_exclusive_scan(sum, sh, excl);
    int run = (int)(offs[blockIdx.x] + excl);
    int o[SCAN_ITEMS];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        o[i] = run;
        run += c[i];
    }
    if (base + SCAN_ITEMS <= n) {
        int4* q = reinterpret_cast<int4*>(out + base);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS / 4; ++k) q[k] = make_int4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            if (base + i < n) out[base + i] = o[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BLOCK - 1) out[n] = (int)total_out;
}

// out (16-byte aligned, n + 1 entries) = exclusive prefix of the byte counts
// of cells 0..n-1 (laid out as count_word / count_shift, padded to whole
// 256-cell blocks), then the total.
static int exclusive_scan_u8(const uint32_t* in, int64_t n, int* out, DevBuf& tmp, int64_t total, hipStream_t s) {
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) nb = 1;
    if (tmp.ensure(sizeof(int64_t) * (size_t)nb)) return PPMG_ERROR_FATAL;
    int64_t* sums = tmp.as<int64_t>();
    hipLaunchKernelGGL(k_scan_reduce_u8, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SUMS_THREADS), 0, s, sums, nb);
    hipLaunchKernelGGL(k_scan_down_u8, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, sums, out, total);
    PPMG_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ binning ---
// Heads of the cells lo..hi (all empty but possibly the last) := v.  Short
// runs are written by the lane itself; longer ones are queued for
// k_fill_gaps in pieces of GAP_CHUNK cells (err[2] counts the entries; an
// overflowing queue sends the caller to the counting path).
constexpr int GAP_LOCAL = 32;
constexpr int GAP_CHUNK = 65536;

__device__ __forceinline__ void set_heads(int* __restrict__ start, int lo, int hi, int v, int* __restrict__ err,
                                          int4* __restrict__ gaps, int gapcap) {
    if (hi - lo < GAP_LOCAL) {
        for (int c = lo; c <= hi; ++c) start[c] = v;
        return;
    }
    for (int64_t a = lo; a <= hi; a += GAP_CHUNK) {
        const int e = (int)(a + GAP_CHUNK - 1 < hi ? a + GAP_CHUNK - 1 : hi);
        const int i = atomicAdd(&err[2], 1);
        if (i < gapcap) gaps[i] = make_int4((int)a, e, v, 0);
    }
}

// Keys of all particles.  err[0] counts the particles in no cell; err[1] is
// set to 1 when a descent key[p] < key[p-1] exists (none: the input is already
// in cell order, e.g. particles made by create_part or a lattice).  A
// wavefront takes KEYS_PPT runs of 64 consecutive particles: all their
// positions are loaded first, then (SLOT) all their atomics issued, so each
// lane keeps KEYS_PPT loads and atomics in flight instead of waiting for one
// round trip at a time (one particle per thread streams the positions at
// 2.9 TB/s, bound by the atomic's latency after each load).  Lanes get their
// predecessor's key by a shuffle, lane 0 from the run before (lane 63) or,
// for the first run, from the previous particle's position, so descents
// across runs and wavefront seams are caught here too.  The flag is read
// before it is written, so the many wavefronts of unsorted input do not
// queue on one word.
//   SLOT (unsorted input expected): write the keys and take each particle's
//     arrival slot in its cell by an atomic count (byte-wide: slots 0..254).
//   SLOT and TABLE: no keys or slots written; particle p, arriving r-th in
//     cell k, is written to table[r * ncells + k] (r < TABLE_SLOTS; a later
//     arrival sets err[4]).
//   !SLOT (input expected in cell order): the first particle of each run of
//     equal keys writes its index as the head of its cell and of the empty
//     cells before it (back to its predecessor's cell); the last particle
//     closes the array with np.  No counts, no scan, no keys written, and the
//     permutation is the identity.  Valid only if no descent is found (else
//     the caller bins again on the SLOT path).
constexpr int KEYS_PPT = 4;                    // runs of 64 particles per wavefront
constexpr int KEYS_BLOCK = 256 * KEYS_PPT;     // particles per key workgroup (and list capacity)

template <int DIM, typename PosT, bool SLOT, bool TABLE>
__global__ void __launch_bounds__(256) k_keys(const PosT* __restrict__ xp, int64_t np, DevMesh m, int64_t ncells,
                                              int* __restrict__ key, uint8_t* __restrict__ slot,
                                              uint32_t* __restrict__ count, int* __restrict__ start,
                                              int* __restrict__ err, int4* __restrict__ gaps, int gapcap,
                                              int* __restrict__ multi, int* __restrict__ multi_n,
                                              int* __restrict__ table) {
    static_assert(SLOT || !TABLE, "the slot table needs the arrival slots");
    const int lane = threadIdx.x & 63;
    // first particle of this wavefront's runs (wavefront-uniform)
    const int64_t w0 = ((int64_t)blockIdx.x * 256 + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u)) * KEYS_PPT;
    const bool has_seam = w0 > 0 && w0 <= np;
    // all loads first (the key arithmetic branches on mesh flags, which would
    // otherwise keep each load behind the previous component's arithmetic)
    // (past the end a lane reads the last particle again, unconditionally:
    // no branch between the loads)
    PosT xv[KEYS_PPT][DIM], sv[DIM];
#pragma unroll
    for (int j = 0; j < KEYS_PPT; ++j) {
        const int64_t p = min(w0 + j * 64 + lane, np - 1);
#pragma unroll
        for (int d = 0; d < DIM; ++d) xv[j][d] = xp[p * DIM + d];
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) sv[d] = has_seam ? xp[(w0 - 1) * DIM + d] : PosT(0);
    int k[KEYS_PPT];
    int nbad = 0;
    uint64_t anydesc = 0, notlat = 0;
    bool desc[KEYS_PPT];
    int prev = has_seam ? cell_key<DIM, PosT>(sv, m) : -1;   // key before run j's lane 0
#pragma unroll
    for (int j = 0; j < KEYS_PPT; ++j) {
        const int64_t p = w0 + j * 64 + lane;
        const bool live = p < np;
        k[j] = live ? cell_key<DIM, PosT>(xv[j], m) : -1;
        int left = __shfl_up(k[j], 1);
        if (lane == 0) left = prev;
        prev = __builtin_amdgcn_readlane(k[j], 63);
        desc[j] = live && p > 0 && k[j] < left;
        nbad += (int)__popcll(__ballot(live && k[j] < 0));
        anydesc |= __ballot(desc[j]);
        notlat |= __ballot(live && k[j] != (int)p);
        if (!SLOT) {
            // heads: the first particle of a run of equal keys
            if (live && k[j] >= 0 && !desc[j]) {
                if (k[j] != left) set_heads(start, left + 1, k[j], (int)p, err, gaps, gapcap);
                if (p == np - 1) set_heads(start, k[j] + 1, (int)ncells, (int)np, err, gaps, gapcap);
            }
        }
    }
    // (last on the SLOT path: the flag's read waits for every memory
    // operation issued before it)
    auto report = [&]() {
        if (lane == 0) {
            if (nbad) atomicAdd(&err[0], nbad);
            if (anydesc && *(volatile int*)&err[1] == 0) err[1] = 1;
            if (notlat && *(volatile int*)&err[5] == 0) err[5] = 1;
        }
    };
    if (!SLOT) report();
    if constexpr (SLOT) {
        // arrival slots, all runs' atomics in flight together: issued without
        // branches (a lane with no cell adds 0 to word 0) and decoded after
        // the last one, so no atomic waits for the one before
        uint32_t raw[KEYS_PPT];
#pragma unroll
        for (int j = 0; j < KEYS_PPT; ++j) {
            // byte-wide counts (count_word / count_shift)
            const bool in = k[j] >= 0;
            raw[j] = atomicAdd(&count[in ? count_word(k[j]) : 0], in ? 1u << count_shift(k[j]) : 0u);
        }
        int sl[KEYS_PPT];
#pragma unroll
        for (int j = 0; j < KEYS_PPT; ++j) sl[j] = k[j] >= 0 ? (int)((raw[j] >> count_shift(k[j])) & 255u) : -1;
        // The particle that arrives second in a cell lists the cell in its
        // workgroup's block of `multi` for the stabilisation.
        __shared__ int listed[KEYS_BLOCK];
        __shared__ int nlisted, ncrowd;
        if (threadIdx.x == 0) nlisted = ncrowd = 0;
        bool ov = false, full = false;
        int crowd = 0;
#pragma unroll
        for (int j = 0; j < KEYS_PPT; ++j) {
            const int64_t p = w0 + j * 64 + lane;
            if (p < np) {
                if (!TABLE) {
                    key[p] = k[j];
                    if (k[j] >= 0) slot[p] = (uint8_t)sl[j];
                } else if (sl[j] >= 0 && sl[j] < TABLE_SLOTS) {
                    table[(int64_t)sl[j] * ncells + k[j]] = (int)p;
                }
            }
            // a cell's 256th arrival carries into the next cell's byte: err[3]
            // sends the caller to the full-width count
            ov |= sl[j] == 255;
            full |= sl[j] >= TABLE_SLOTS;
            crowd += sl[j] == 8;
        }
        if (__ballot(ov) && lane == 0) err[3] = 1;
        if (TABLE && __ballot(full) && lane == 0 && *(volatile int*)&err[4] == 0) err[4] = 1;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KEYS_PPT; ++j)
            if (sl[j] == 1) listed[atomicAdd(&nlisted, 1)] = k[j];
        if ((blockIdx.x & 63) == 0 && crowd > 0) atomicAdd(&ncrowd, crowd);
        __syncthreads();
        const int n = nlisted;
        // err[8 .. 71]: the cells whose ninth particle arrived in one key
        // workgroup of every 64 (a sample, for the deposit's choice of chunk),
        // spread over 64 words: contended device-scope atomics serialise (one
        // per crowded cell took 27 ms at 512^3 clustered, one per sampled
        // workgroup on a single word 0.85 ms)
        if (threadIdx.x == 0 && ncrowd > 0 && (blockIdx.x & 63) == 0)
            atomicAdd(&err[8 + ((blockIdx.x >> 6) & 63)], ncrowd);
        for (int i = (int)threadIdx.x; i < n; i += 256) multi[(int64_t)blockIdx.x * KEYS_BLOCK + i] = listed[i];
        if (threadIdx.x == 0) multi_n[blockIdx.x] = n;
        report();
    }
}

// Long runs of empty cells queued by k_keys: one workgroup per entry.
__global__ void __launch_bounds__(256) k_fill_gaps(const int4* __restrict__ gaps, int n, int* __restrict__ start) {
    for (int g = blockIdx.x; g < n; g += gridDim.x) {
        const int4 e = gaps[g];
        for (int c = e.x + (int)threadIdx.x; c <= e.y; c += 256) start[c] = e.z;
    }
}

// Full-width arrival slots of every particle in its cell (a cell with more
// particles than a byte counts).
__global__ void __launch_bounds__(256) k_count(const int* __restrict__ key, int64_t np, int* __restrict__ slot,
                                               int* __restrict__ count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    slot[p] = atomicAdd(&count[key[p]], 1);
}

__global__ void __launch_bounds__(256) k_iota(int* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int)i;
}

__global__ void __launch_bounds__(256) k_identity(int64_t np, int* __restrict__ perm) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) perm[p] = (int)p;
}


// Cell key with the arithmetic of ppm_util_rank3d.f:414-475 (FLOOR, ghost
// cells, upper-boundary corrections, -1 for particles in no box).
template <int DIM, typename PosT>
__global__ void __launch_bounds__(256) k_rank_keys(const PosT* __restrict__ xp, int64_t np,
                                                   RankGeom g, int* __restrict__ key,
                                                   int* __restrict__ slot, int* __restrict__ count,
                                                   int* __restrict__ stats) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const PosT* x = xp + p * DIM;
    int idx[3] = {0, 0, 0};
    int corr = 0;
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        PosT mn = (PosT)g.xmin[d], mx = (PosT)g.xmax[d];
        PosT rd = (PosT)g.rdx[d];
        int i = (int)floor(scaled_coord<PosT>(x[d], mn, rd)) + g.ngl_lo[d];
        if (x[d] >= mx && i < g.nm[d] + g.ngl_lo[d]) { i = g.nm[d] + g.ngl_lo[d]; ++corr; }
        idx[d] = i;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        PosT mx = (PosT)g.xmax[d];
        if (x[d] < mx && idx[d] >= g.nm[d] + g.ngl_lo[d]) { idx[d] = g.nm[d] + g.ngl_lo[d] - 1; ++corr; }
        if (idx[d] < 0 || idx[d] >= g.ntot[d]) in = false;
    }
    if (corr) atomicAdd(&stats[1], corr);
    if (!in) {
        key[p] = -1;
        atomicAdd(&stats[0], 1);
        return;
    }
    int k = idx[0] + g.ntot[0] * (idx[1] + g.ntot[1] * idx[2]);
    key[p] = k;
    slot[p] = atomicAdd(&count[k], 1);
}

// Four particles per thread: keys and slots as vector loads (plan buffers
// are 256-byte aligned), then the four head loads in flight together.  SlotT:
// uint8_t (byte-wide counting) or int (full width).
template <typename SlotT>
__global__ void __launch_bounds__(256) k_scatter(const int* __restrict__ key, const SlotT* __restrict__ slot,
                                                 const int* __restrict__ start, int64_t np,
                                                 int* __restrict__ perm) {
    const int64_t p0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (p0 >= np) return;
    int k[4], sl[4];
    if (p0 + 4 <= np) {
        const int4 kv = *reinterpret_cast<const int4*>(key + p0);
        k[0] = kv.x; k[1] = kv.y; k[2] = kv.z; k[3] = kv.w;
        if constexpr (sizeof(SlotT) == 1) {
            const uint32_t sv = *reinterpret_cast<const uint32_t*>(slot + p0);
#pragma unroll
            for (int i = 0; i < 4; ++i) sl[i] = (int)((sv >> (8 * i)) & 255u);
        } else {
            const int4 sv = *reinterpret_cast<const int4*>(slot + p0);
            sl[0] = sv.x; sl[1] = sv.y; sl[2] = sv.z; sl[3] = sv.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            k[i] = p0 + i < np ? key[p0 + i] : -1;
            sl[i] = p0 + i < np ? (int)slot[p0 + i] : 0;
        }
    }
    int h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = k[i] >= 0 ? start[k[i]] : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (k[i] >= 0) perm[h[i] + sl[i]] = (int)(p0 + i);
}

// Restore increasing particle index inside every cell (stability).
__global__ void __launch_bounds__(256) k_stabilize(const int* __restrict__ start, int64_t ncells,
                                                   int* __restrict__ perm) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int b = start[c], e = start[c + 1];
    if (e - b < 2) return;
    for (int i = b + 1; i < e; ++i) {
        int v = perm[i];
        int j = i - 1;
        while (j >= b && perm[j] > v) {
            perm[j + 1] = perm[j];
            --j;
        }
        perm[j + 1] = v;
    }
}

// The same for the cells k_keys listed (those with two or more particles).
// A key workgroup lists a few hundred cells at most (about 280 of its 1024
// particles on an advected lattice), so one wavefront takes the lists of
// STAB_GROUP consecutive key workgroups, spread over its lanes.
constexpr int STAB_GROUP = 2;
__global__ void __launch_bounds__(64) k_stabilize_listed(const int* __restrict__ start,
                                                         const int* __restrict__ multi,
                                                         const int* __restrict__ multi_n, int64_t nlists,
                                                         int* __restrict__ perm) {
    const int64_t l0 = (int64_t)blockIdx.x * STAB_GROUP;
    int off[STAB_GROUP + 1];
    off[0] = 0;
#pragma unroll
    for (int j = 0; j < STAB_GROUP; ++j) off[j + 1] = off[j] + (l0 + j < nlists ? multi_n[l0 + j] : 0);
    for (int q = (int)threadIdx.x; q < off[STAB_GROUP]; q += 64) {
        int j = 0;
        while (q >= off[j + 1]) ++j;
        const int c = multi[(l0 + j) * KEYS_BLOCK + (q - off[j])];
        const int b = start[c], e = start[c + 1];
        for (int i = b + 1; i < e; ++i) {
            const int v = perm[i];
            int k = i - 1;
            while (k >= b && perm[k] > v) {
                perm[k + 1] = perm[k];
                --k;
            }
            perm[k + 1] = v;
        }
    }
}

// The same for the listed cells of a slot table: the (at most TABLE_SLOTS)
// entries of a cell are sorted in registers.
__global__ void __launch_bounds__(64) k_stabilize_table(const uint32_t* __restrict__ count,
                                                        const int* __restrict__ multi,
                                                        const int* __restrict__ multi_n, int64_t nlists,
                                                        int64_t ncells, int* __restrict__ table) {
    const int64_t l0 = (int64_t)blockIdx.x * STAB_GROUP;
    int off[STAB_GROUP + 1];
    off[0] = 0;
#pragma unroll
    for (int j = 0; j < STAB_GROUP; ++j) off[j + 1] = off[j] + (l0 + j < nlists ? multi_n[l0 + j] : 0);
    for (int q = (int)threadIdx.x; q < off[STAB_GROUP]; q += 64) {
        int j = 0;
        while (q >= off[j + 1]) ++j;
        const int c = multi[(l0 + j) * KEYS_BLOCK + (q - off[j])];
        const int n = (int)((count[count_word(c)] >> count_shift(c)) & 255u);
        int v[TABLE_SLOTS];
#pragma unroll
        for (int r = 0; r < TABLE_SLOTS; ++r) v[r] = r < n ? table[(int64_t)r * ncells + c] : INT_MAX;
        // odd-even transposition sort (the padding INT_MAX stays at the end)
#pragma unroll
        for (int pass = 0; pass < TABLE_SLOTS; ++pass)
#pragma unroll
            for (int r = pass & 1; r + 1 < TABLE_SLOTS; r += 2) {
                const int a = min(v[r], v[r + 1]), b = max(v[r], v[r + 1]);
                v[r] = a;
                v[r + 1] = b;
            }
#pragma unroll
        for (int r = 0; r < TABLE_SLOTS; ++r)
            if (r < n) table[(int64_t)r * ncells + c] = v[r];
    }
}

// Permutation of a slot-table binning from its heads: the particles of cell c
// go to perm[start[c] ..].
__global__ void __launch_bounds__(256) k_table_perm(const uint32_t* __restrict__ count, const int* __restrict__ start,
                                                    const int* __restrict__ table, int64_t ncells,
                                                    int* __restrict__ perm) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int n = (int)((count[count_word(c)] >> count_shift(c)) & 255u);
    const int b = start[c];
    for (int r = 0; r < n; ++r) perm[b + r] = table[(int64_t)r * ncells + c];
}

__global__ void k_to_one_based(int* __restrict__ a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += 1;
}

__global__ void k_fill_zero(uint32_t* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

// Zero fill with a kernel (graph-capture safe, unlike hipMemsetAsync).
int fill_zero(void* p, size_t bytes, hipStream_t s) {
    int64_t n = (int64_t)(bytes / 4);
    if (n <= 0) return 0;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)g), dim3(256), 0, s, (uint32_t*)p, n);
    PPMG_HIP(hipGetLastError());
    return 0;
}

// Keys of a binning that took the cell-order path (which writes none).
template <int DIM, typename PosT>
__global__ void __launch_bounds__(256) k_keys_only(const PosT* __restrict__ xp, int64_t np, DevMesh m,
                                                   int* __restrict__ key) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    PosT xv[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xv[d] = xp[p * DIM + d];
    key[p] = cell_key<DIM, PosT>(xv, m);
}

// ------------------------------------------------------ crowded input ---
// Cells of tens of particles in no particular order (clustered particles):
// k_keys' atomic slots contend on the crowded cells' count words, the scatter
// writes at random and the per-cell insertion sorts grow with the square of
// the occupancy.  Such input is binned by a stable LSD radix sort of the
// (cell key, particle) pairs instead, RS_BITS bits a pass: no atomics on
// crowded words, writes in runs, and a stable sort of the keys in particle
// order is the reference's permutation as it stands (ppm_util_rank3d.f:
// 505-532).  The heads follow from the runs of the sorted keys.
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;   // pairs per workgroup
constexpr int RS_BITS = 8;
constexpr int RS_BUCKETS = 1 << RS_BITS;
static_assert(RS_BUCKETS == RS_THREADS, "one digit per thread in the tile loops");

// Keys of all particles (key[p]), the particles in no cell (err[0]) and the
// descent flag (err[1]), as k_keys reports them.
template <int DIM, typename PosT>
__global__ void __launch_bounds__(256) k_keys_plain(const PosT* __restrict__ xp, int64_t np, DevMesh m,
                                                    int* __restrict__ key, int* __restrict__ err) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = p < np;
    const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    const bool has_seam = p0 > 0 && p0 <= np;
    PosT xv[DIM], sv[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xv[d] = live ? xp[p * DIM + d] : PosT(0);
        sv[d] = has_seam ? xp[(p0 - 1) * DIM + d] : PosT(0);
    }
    const int seam = has_seam ? cell_key<DIM, PosT>(sv, m) : -1;
    const int k = live ? cell_key<DIM, PosT>(xv, m) : -1;
    int left = __shfl_up(k, 1);
    if (lane == 0) left = seam;
    const uint64_t bb = __ballot(live && k < 0), bd = __ballot(live && p > 0 && k < left);
    if (live) key[p] = k;
    if (lane == 0) {
        if (bb) atomicAdd(&err[0], (int)__popcll(bb));
        if (bd && *(volatile int*)&err[1] == 0) err[1] = 1;
    }
}

// Lanes of the wavefront holding the same digit as this lane (among `live`).
__device__ __forceinline__ uint64_t digit_peers(int d, bool live) {
    uint64_t peers = __ballot(live);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
        const uint64_t on = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? on : ~on;
    }
    return peers;
}

// Digit counts of every tile, digit-major: hist[d * ntiles + tile].  Tiles
// are numbered XCD-contiguously (xcd_tile), as in k_rs_scatter: the tiles
// one XCD sorts are neighbours, and so are their runs of each digit.
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const int* __restrict__ keys, int64_t n, int shift,
                                                        int* __restrict__ hist, int64_t ntiles) {
    __shared__ int h[RS_BUCKETS];
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    h[tid] = 0;
    __syncthreads();
    const int64_t t0 = tile * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t i = t0 + j * RS_THREADS + tid;
        const bool live = i < n;
        const int d = live ? (keys[i] >> shift) & (RS_BUCKETS - 1) : 0;
        // one LDS add per digit and wavefront
        const uint64_t peers = digit_peers(d, live);
        if (live && (int)__builtin_ctzll(peers) == lane) atomicAdd(&h[d], (int)__popcll(peers));
    }
    __syncthreads();
    hist[(int64_t)tid * ntiles + tile] = h[tid];
}

// One stable pass.  The pairs of a tile are first ordered by digit in LDS: a
// tile is taken RS_THREADS pairs a round in index order, and a pair's place
// is its digit's start in the tile plus its rank among the tile's pairs of
// that digit (the count from earlier rounds, earlier wavefronts of its round
// and earlier lanes of its wavefront).  The tile then writes its digit runs
// at their offsets (offs: the exclusive scan of hist), consecutive threads to
// consecutive addresses.  vin nullptr: the values are the pairs' indices
// (the first pass).
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const int* __restrict__ kin, const int* __restrict__ vin,
                                                           int64_t n, int shift, const int* __restrict__ hist,
                                                           const int* __restrict__ offs, int64_t ntiles,
                                                           int* __restrict__ kout, int* __restrict__ vout) {
    constexpr int NWAVE = RS_THREADS / 64;
    __shared__ int base[RS_BUCKETS];     // next place of each digit in the tile's order
    __shared__ int shift_out[RS_BUCKETS];   // global position minus place, per digit
    __shared__ int wcnt[NWAVE][RS_BUCKETS];
    __shared__ int sk[RS_TILE], sv[RS_TILE];
    __shared__ int64_t sh[RS_THREADS / 64 + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t tile = xcd_tile(blockIdx.x, (int64_t)gridDim.x);
    const int64_t t0 = tile * RS_TILE;
    const int tn = (int)min((int64_t)RS_TILE, n - t0);
    // thread t = digit t: its start in the tile (exclusive scan of the tile's counts)
    int64_t lb;
    block_exclusive_scan((int64_t)hist[(int64_t)tid * ntiles + tile], sh, lb);
    base[tid] = (int)lb;
    shift_out[tid] = offs[(int64_t)tid * ntiles + tile] - (int)lb;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) wcnt[w][tid] = 0;
    const uint64_t lt = lane ? ~0ull >> (64 - lane) : 0ull;   // lanes below this one
    __syncthreads();
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int64_t i = t0 + j * RS_THREADS + tid;
        const bool live = i < n;
        const int k = live ? kin[i] : 0;
        const int v = live ? (vin ? vin[i] : (int)i) : 0;
        const int d = (k >> shift) & (RS_BUCKETS - 1);
        const uint64_t peers = digit_peers(d, live);
        if (live && (int)__builtin_ctzll(peers) == lane) wcnt[wave][d] = (int)__popcll(peers);
        __syncthreads();
        if (live) {
            int r = base[d] + (int)__popcll(peers & lt);
            for (int w = 0; w < wave; ++w) r += wcnt[w][d];
            sk[r] = k;
            sv[r] = v;
        }
        __syncthreads();
        // thread t owns digit t: advance its place, clear its counts
        int add = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        base[tid] += add;
        __syncthreads();
    }
    for (int q = tid; q < tn; q += RS_THREADS) {
        const int k = sk[q];
        const int o = shift_out[(k >> shift) & (RS_BUCKETS - 1)] + q;
        kout[o] = k;
        vout[o] = sv[q];
    }
}

// Cell counts from the sorted keys: each run of equal keys adds its length
// to cnt[key] (zeroed) by two atomics at its ends, on distinct runs' words;
// a run of one particle stores its count.
// (Writing the heads straight from the runs, as the cell-order path does,
// took longer on crowded input: most cells are empty, and a lane fills each
// run of empty cells before its own or queues it on one counter word.)  One
// workgroup in 64 counts its runs of nine or more into err[8 .. 71] (the
// same sample k_keys takes).
__global__ void __launch_bounds__(256) k_rs_runs(const int* __restrict__ ks, int64_t n, int* __restrict__ cnt,
                                                 int* __restrict__ err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool crowd = false;
    if (i < n) {
        const int k = ks[i];
        const bool first = i == 0 || ks[i - 1] != k, last = i == n - 1 || ks[i + 1] != k;
        if (first && last) {
            cnt[k] = 1;
        } else {
            if (first) atomicAdd(&cnt[k], -(int)i);
            if (last) atomicAdd(&cnt[k], (int)i + 1);
        }
        crowd = last && i >= 8 && ks[i - 8] == k;
    }
    if ((blockIdx.x & 63) == 0) {
        const uint64_t c = __ballot(crowd);
        if (c && (threadIdx.x & 63) == 0) atomicAdd(&err[8 + ((blockIdx.x >> 6) & 63)], (int)__popcll(c));
    }
}

// Radix-sort binning of crowded input (see above); the keys pass has run.
static int bin_radix(ppmg_plan* plan, int64_t np, int64_t ncells, hipStream_t s, const int** sorted_keys) {
    const int64_t ntiles = (np + RS_TILE - 1) / RS_TILE;
    if (plan->rkey[0].ensure(sizeof(int) * (size_t)np) || plan->rkey[1].ensure(sizeof(int) * (size_t)np) ||
        plan->rval[0].ensure(sizeof(int) * (size_t)np) || plan->rval[1].ensure(sizeof(int) * (size_t)np) ||
        plan->rhist.ensure(sizeof(int) * (size_t)(2 * (RS_BUCKETS * ntiles + 4))))
        return PPMG_ERROR_FATAL;
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < ncells) ++bits;
    const int passes = (bits + RS_BITS - 1) / RS_BITS;
    int* hist = plan->rhist.as<int>();
    int* offs = hist + RS_BUCKETS * ntiles + 4;   // 16-byte aligned, n + 1 entries
    const int* kin = plan->key.as<int>();
    const int* vin = nullptr;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = pass * RS_BITS;
        const bool last = pass == passes - 1;
        int* kout = plan->rkey[pass & 1].as<int>();
        int* vout = last ? plan->perm.as<int>() : plan->rval[pass & 1].as<int>();
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, kin, np, shift, hist, ntiles);
        if (exclusive_scan(hist, RS_BUCKETS * ntiles, offs, plan->scan_tmp, np, s)) return PPMG_ERROR_FATAL;
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, kin, vin, np, shift, hist,
                           offs, ntiles, kout, vout);
        kin = kout;
        vin = vout;
    }
    *sorted_keys = kin;
    PPMG_HIP(hipGetLastError());
    return 0;
}

// Heads of the sorted keys ks (run lengths into zeroed counts, then their
// exclusive scan), and the plan's crowded-cell estimate from the sample.
static int heads_from_sorted(ppmg_plan* plan, const int* ks, int64_t np, int64_t ncells, hipStream_t s) {
    int* err = plan->err.as<int>();
    if (plan->count.ensure(sizeof(int) * (size_t)ncells)) return PPMG_ERROR_FATAL;
    int* cnt = plan->count.as<int>();
    if (fill_zero(cnt, sizeof(int) * (size_t)ncells, s)) return PPMG_ERROR_FATAL;
    hipLaunchKernelGGL(k_rs_runs, dim3(nblk(np, 256)), dim3(256), 0, s, ks, np, cnt, err);
    if (exclusive_scan(cnt, ncells, plan->start.as<int>(), plan->scan_tmp, np, s)) return PPMG_ERROR_FATAL;
    PPMG_HIP(hipGetLastError());
    int sample[64];
    PPMG_HIP(hipMemcpyAsync(sample, err + 8, sizeof(sample), hipMemcpyDeviceToHost, s));
    PPMG_HIP(hipStreamSynchronize(s));
    int64_t sampled = 0;
    for (int v : sample) sampled += v;
    plan->crowded_cells = 64 * sampled;
    return 0;
}

// The binning of crowded input: keys pass, then (input not in cell order)
// the radix sort, and the heads; the plan's crowded-cell estimate is taken
// again from the sorted runs.
static int bin_crowded(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp, hipStream_t s,
                       int64_t* nbad) {
    const int64_t ncells = plan->ncells;
    int* key = plan->key.as<int>();
    int* err = plan->err.as<int>();
    const unsigned g = nblk(np, 256);
    if (m.dim == 3) {
        if (prec == PPMG_PREC_F32) hipLaunchKernelGGL((k_keys_plain<3, float>), dim3(g), dim3(256), 0, s, (const float*)xp, np, m, key, err);
        else hipLaunchKernelGGL((k_keys_plain<3, double>), dim3(g), dim3(256), 0, s, (const double*)xp, np, m, key, err);
    } else {
        if (prec == PPMG_PREC_F32) hipLaunchKernelGGL((k_keys_plain<2, float>), dim3(g), dim3(256), 0, s, (const float*)xp, np, m, key, err);
        else hipLaunchKernelGGL((k_keys_plain<2, double>), dim3(g), dim3(256), 0, s, (const double*)xp, np, m, key, err);
    }
    PPMG_HIP(hipGetLastError());
    int herr[2] = {0, 0};
    PPMG_HIP(hipMemcpyAsync(herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
    PPMG_HIP(hipStreamSynchronize(s));
    *nbad = herr[0];
    if (herr[0]) return 0;
    // input already in cell order: the keys are the sorted keys, the
    // permutation the identity (written on demand)
    const bool sorted = herr[1] == 0;
    const int* ks = key;
    if (!sorted && bin_radix(plan, np, ncells, s, &ks)) return PPMG_ERROR_FATAL;
    if (heads_from_sorted(plan, ks, np, ncells, s)) return PPMG_ERROR_FATAL;
    plan->perm_identity = sorted;
    plan->perm_written = !sorted;
    plan->keys_written = true;
    plan->sorted_input = sorted;
    return 0;
}

int bin_particles(ppmg_plan* plan, const DevMesh& m, int prec, int64_t np, const void* xp,
                  hipStream_t s, int64_t* nbad, bool table, bool lattice) {
    const int64_t ncells = (int64_t)m.nbin[0] * m.nbin[1] * m.nbin[2];
    const int64_t gapcap = np / 16 + 4096;
    const size_t npb = (size_t)(np > 0 ? np : 1);
    // byte-wide counts (count_word) come in whole blocks of 256 cells, and
    // the scan reads whole tiles of SCAN_TILE cells
    const size_t count_alloc = ((size_t)ncells + SCAN_TILE - 1) / SCAN_TILE * SCAN_TILE;
    if (plan->key.ensure(sizeof(int) * npb) || plan->slot.ensure(npb + 16) ||
        plan->perm.ensure(sizeof(int) * npb) || plan->count.ensure(count_alloc) ||
        plan->start.ensure(sizeof(int) * (size_t)(ncells + 1)) ||
        plan->gaps.ensure(sizeof(int4) * (size_t)gapcap) ||
        plan->multi.ensure(sizeof(int) * (size_t)nblk(np, KEYS_BLOCK) * KEYS_BLOCK) ||
        plan->multi_n.ensure(sizeof(int) * (size_t)nblk(np, KEYS_BLOCK)) ||
        plan->err.ensure(sizeof(int) * ERR_WORDS))
        return PPMG_ERROR_FATAL;
    plan->bin_stream = s;
    plan->prebinned = false;   // a binning left by ppmg_m2p_prebin is overwritten
    plan->bin_table = false;
    plan->bin_lattice = false;
    plan->lattice_checked = false;
    plan->np = np;
    plan->ncells = ncells;
    plan->bin_xp = xp;
    plan->bin_prec = prec;
    plan->bin_mesh = m;
    plan->keys_written = true;
    *nbad = 0;
    if (np == 0) {
        plan->perm_identity = true;
        plan->perm_written = true;
        return fill_zero(plan->start.p, sizeof(int) * (size_t)(ncells + 1), s);
    }
    if (fill_zero(plan->err.p, sizeof(int) * ERR_WORDS, s)) return PPMG_ERROR_FATAL;
    // A lattice in cell order (particle p in cell p, np == ncells), as the last
    // binning of this plan found on the same mesh: for a reader that checks
    // every particle's cell itself (`lattice`), no pass at all; the reader
    // reports a particle outside its cell in err[6] and bins again.
    // A plan whose lattice check failed takes its next 16 inputs without it.
    if (lattice && plan->lattice_skip > 0) {
        --plan->lattice_skip;
        lattice = false;
    }
    if (lattice && plan->lattice_pred && np == ncells && plan->lattice_prec == prec &&
        same_mesh(plan->lattice_mesh, m)) {
        plan->bin_lattice = true;
        plan->perm_identity = true;
        plan->perm_written = false;
        plan->keys_written = false;
        return 0;
    }
    plan->lattice_pred = false;
    // The plan remembers whether its last input was in cell order; a plan fed
    // the same kind of particles each call (advected particles, or particles
    // fresh from create_part) predicts right and does each step once.
    const bool predict_sorted = plan->sorted_input;
    // Unsorted input for a reader of the slot table: the key pass fills the
    // table, and no keys, scan or scatter follow.  A plan whose table
    // overflowed bins its next 16 inputs without one.  Allocation-free calls
    // take the heads (the table needs 32 bytes per cell), and so does every
    // call under PPMG_NO_TABLE=1 (used to cross-check the two layouts).
    const char* no_table = std::getenv("PPMG_NO_TABLE");
    if (no_table && no_table[0] && no_table[0] != '0') table = false;
    const bool use_table = table && !predict_sorted && plan->table_skip == 0 && !arena_bound();
    if (table && !predict_sorted && plan->table_skip > 0) --plan->table_skip;
    // Crowded input (the last binning's sample found a cell of nine or more
    // particles per 64 particles) is binned by the radix sort; PPMG_RADIX=1
    // forces that path for input not predicted in cell order, PPMG_RADIX=0
    // rules it out (both for tests).  Allocation-free calls keep the counting
    // path (the sort needs four more int32 per particle).
    const char* rx = std::getenv("PPMG_RADIX");
    const bool radix = rx && rx[0] ? rx[0] == '1' : plan->crowded_cells * 64 > np;
    if (!predict_sorted && radix && !arena_bound()) return bin_crowded(plan, m, prec, np, xp, s, nbad);
    if (use_table && plan->table.ensure(sizeof(int) * TABLE_SLOTS * (size_t)ncells)) return PPMG_ERROR_FATAL;
    int* tab = use_table ? plan->table.as<int>() : nullptr;
    if (!predict_sorted && fill_zero(plan->count.p, count_alloc, s)) return PPMG_ERROR_FATAL;
    int* key = plan->key.as<int>();
    uint8_t* slot8 = plan->slot.as<uint8_t>();
    uint32_t* count8 = plan->count.as<uint32_t>();
    int* start = plan->start.as<int>();
    int* err = plan->err.as<int>();
    int4* gaps = plan->gaps.as<int4>();
    int* multi = plan->multi.as<int>();
    int* multi_n = plan->multi_n.as<int>();
    {
        const unsigned g = nblk(np, KEYS_BLOCK);
#define KEYS(D, T, S, TB) hipLaunchKernelGGL((k_keys<D, T, S, TB>), dim3(g), dim3(256), 0, s, (const T*)xp, np, m, \
                                             ncells, key, slot8, count8, start, err, gaps, (int)gapcap, multi, multi_n, tab)
#define KEYS_T(D, T)                                   \
    do {                                               \
        if (predict_sorted) KEYS(D, T, false, false);  \
        else if (use_table) KEYS(D, T, true, true);    \
        else KEYS(D, T, true, false);                  \
    } while (0)
        const bool single = prec == PPMG_PREC_F32;
        if (m.dim == 3) {
            if (single) KEYS_T(3, float);
            else KEYS_T(3, double);
        } else {
            if (single) KEYS_T(2, float);
            else KEYS_T(2, double);
        }
#undef KEYS_T
#undef KEYS
        PPMG_HIP(hipGetLastError());
    }
    int herr[ERR_WORDS] = {};
    PPMG_HIP(hipMemcpyAsync(herr, plan->err.p, sizeof(int) * ERR_WORDS, hipMemcpyDeviceToHost, s));
    PPMG_HIP(hipStreamSynchronize(s));
    *nbad = herr[0];
    int64_t sampled = 0;
    for (int i = 8; i < ERR_WORDS; ++i) sampled += herr[i];
    plan->crowded_cells = 64 * sampled;   // estimate from the 1-in-64 sample
    if (herr[0]) return 0;
    plan->lattice_pred = herr[5] == 0 && np == ncells;
    plan->lattice_prec = prec;
    plan->lattice_mesh = m;
    const bool sorted = herr[1] == 0;
    if (predict_sorted) {
        if (sorted && herr[2] <= gapcap) {
            // the heads are written but for the queued runs of empty cells
            if (herr[2] > 0)
                hipLaunchKernelGGL(k_fill_gaps, dim3((unsigned)(herr[2] < 4096 ? herr[2] : 4096)), dim3(256), 0, s,
                                   gaps, herr[2], start);
            PPMG_HIP(hipGetLastError());
            // input in cell order: identity permutation and keys, both written only on demand
            plan->perm_identity = true;
            plan->perm_written = false;
            plan->keys_written = false;
            return 0;
        }
        // wrong prediction (or more long gaps than the queue holds): bin
        // again on the counting path
        plan->sorted_input = false;
        return bin_particles(plan, m, prec, np, xp, s, nbad, table, false);
    }
    if (use_table) {
        if (herr[4]) {
            // a cell holds more than TABLE_SLOTS particles: bin again with
            // keys and slots
            plan->table_skip = 16;
            return bin_particles(plan, m, prec, np, xp, s, nbad, false, false);
        }
        // the listed cells (two or more particles) in increasing particle index
        hipLaunchKernelGGL(k_stabilize_table, dim3(nblk(nblk(np, KEYS_BLOCK), STAB_GROUP)), dim3(64), 0, s, count8,
                           multi, multi_n, (int64_t)nblk(np, KEYS_BLOCK), ncells, tab);
        PPMG_HIP(hipGetLastError());
        plan->bin_table = true;
        plan->perm_identity = false;
        plan->perm_written = false;
        plan->keys_written = false;
        plan->sorted_input = herr[1] == 0;
        return 0;
    }
    if (!herr[3]) {
        if (exclusive_scan_u8(count8, ncells, start, plan->scan_tmp, np, s)) return PPMG_ERROR_FATAL;
        if (!sorted)
            hipLaunchKernelGGL(k_scatter<uint8_t>, dim3(nblk(np, 1024)), dim3(256), 0, s, key, slot8, start, np,
                               plan->perm.as<int>());
    } else {
        // a cell holds more particles than a byte counts: count again at full width
        if (plan->slot.ensure(sizeof(int) * npb) || plan->count.ensure(sizeof(int) * (size_t)ncells))
            return PPMG_ERROR_FATAL;
        int* slot32 = plan->slot.as<int>();
        int* count32 = plan->count.as<int>();
        if (fill_zero(count32, sizeof(int) * (size_t)ncells, s)) return PPMG_ERROR_FATAL;
        hipLaunchKernelGGL(k_count, dim3(nblk(np, 256)), dim3(256), 0, s, key, np, slot32, count32);
        if (exclusive_scan(count32, ncells, start, plan->scan_tmp, np, s)) return PPMG_ERROR_FATAL;
        if (!sorted)
            hipLaunchKernelGGL(k_scatter<int>, dim3(nblk(np, 1024)), dim3(256), 0, s, key, slot32, start, np,
                               plan->perm.as<int>());
    }
    if (!sorted) {
        // the key pass listed the cells whose second particle it saw; after a
        // byte overflow that list may miss cells, so every cell is sorted
        if (herr[3])
            hipLaunchKernelGGL(k_stabilize, dim3(nblk(ncells, 256)), dim3(256), 0, s, start, ncells,
                               plan->perm.as<int>());
        else
            hipLaunchKernelGGL(k_stabilize_listed, dim3(nblk(nblk(np, KEYS_BLOCK), STAB_GROUP)), dim3(64), 0, s,
                               start, multi, multi_n, (int64_t)nblk(np, KEYS_BLOCK), plan->perm.as<int>());
    }
    PPMG_HIP(hipGetLastError());
    // input in cell order: the permutation is the identity, written only on demand
    plan->perm_identity = sorted;
    plan->perm_written = !sorted;
    plan->sorted_input = sorted;
    return 0;
}

int materialize_keys(ppmg_plan* plan, hipStream_t s) {
    if (!plan->keys_written && plan->np > 0) {
        const unsigned g = nblk(plan->np, 256);
        const DevMesh& m = plan->bin_mesh;
        int* key = plan->key.as<int>();
        if (plan->bin_prec == PPMG_PREC_F32) {
            if (m.dim == 3) hipLaunchKernelGGL((k_keys_only<3, float>), dim3(g), dim3(256), 0, s, (const float*)plan->bin_xp, plan->np, m, key);
            else hipLaunchKernelGGL((k_keys_only<2, float>), dim3(g), dim3(256), 0, s, (const float*)plan->bin_xp, plan->np, m, key);
        } else {
            if (m.dim == 3) hipLaunchKernelGGL((k_keys_only<3, double>), dim3(g), dim3(256), 0, s, (const double*)plan->bin_xp, plan->np, m, key);
            else hipLaunchKernelGGL((k_keys_only<2, double>), dim3(g), dim3(256), 0, s, (const double*)plan->bin_xp, plan->np, m, key);
        }
        PPMG_HIP(hipGetLastError());
    }
    plan->keys_written = true;
    return 0;
}

int materialize_heads(ppmg_plan* plan, hipStream_t s) {
    if (plan->bin_lattice) {
        if (!plan->lattice_checked) {
            // a lattice binning no gather has checked: bin the positions
            int64_t nbad = 0;
            return bin_particles(plan, plan->bin_mesh, plan->bin_prec, plan->np, plan->bin_xp, s, &nbad, false,
                                 false);
        }
        // a checked lattice: cell c holds particle c
        hipLaunchKernelGGL(k_iota, dim3(nblk(plan->ncells + 1, 256)), dim3(256), 0, s, plan->start.as<int>(),
                           plan->ncells + 1);
        PPMG_HIP(hipGetLastError());
        plan->bin_lattice = false;
        return 0;
    }
    if (!plan->bin_table) return 0;
    const int64_t ncells = plan->ncells, np = plan->np;
    int* start = plan->start.as<int>();
    const uint32_t* count8 = plan->count.as<uint32_t>();
    if (exclusive_scan_u8(count8, ncells, start, plan->scan_tmp, np, s)) return PPMG_ERROR_FATAL;
    hipLaunchKernelGGL(k_table_perm, dim3(nblk(ncells, 256)), dim3(256), 0, s, count8, start,
                       plan->table.as<int>(), ncells, plan->perm.as<int>());
    PPMG_HIP(hipGetLastError());
    plan->bin_table = false;
    plan->perm_identity = false;
    plan->perm_written = true;
    return 0;
}

int materialize_perm(ppmg_plan* plan, hipStream_t s) {
    if (materialize_heads(plan, s)) return PPMG_ERROR_FATAL;
    if (plan->perm_identity && !plan->perm_written && plan->np > 0) {
        hipLaunchKernelGGL(k_identity, dim3(nblk(plan->np, 256)), dim3(256), 0, s, plan->np, plan->perm.as<int>());
        PPMG_HIP(hipGetLastError());
    }
    plan->perm_written = true;
    return 0;
}

// ppm_util_sort3d / sort2d (ppm_util_sort3d.f:237-260): particle rows in the
// ranking's order (1-based lpdx), and the particle count of every box.
template <typename T>
__global__ void __launch_bounds__(256) k_gather_rows(const T* __restrict__ src, const int* __restrict__ lpdx1,
                                                     int64_t n, int dim, T* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = lpdx1[i] - 1;
    for (int d = 0; d < dim; ++d) dst[i * dim + d] = src[p * dim + d];
}

__global__ void __launch_bounds__(256) k_box_counts(const int* __restrict__ lhbx1, int64_t nbox,
                                                    int* __restrict__ npbx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nbox) npbx[c] = lhbx1[c + 1] - lhbx1[c];
}

int sort_rows(int dim, int prec, int64_t np, void* xp, const int* lpdx1, const int* lhbx1, int64_t nbox, int* npbx,
              DevBuf& tmp, hipStream_t s) {
    const size_t esz = prec == PPMG_PREC_F32 ? 4 : 8;
    if (np > 0) {
        if (tmp.ensure(esz * (size_t)np * dim)) return PPMG_ERROR_FATAL;
        if (esz == 4)
            hipLaunchKernelGGL(k_gather_rows<float>, dim3(nblk(np, 256)), dim3(256), 0, s, (const float*)xp, lpdx1, np,
                               dim, tmp.as<float>());
        else
            hipLaunchKernelGGL(k_gather_rows<double>, dim3(nblk(np, 256)), dim3(256), 0, s, (const double*)xp, lpdx1,
                               np, dim, tmp.as<double>());
        PPMG_HIP(hipMemcpyAsync(xp, tmp.p, esz * (size_t)np * dim, hipMemcpyDeviceToDevice, s));
    }
    if (npbx) hipLaunchKernelGGL(k_box_counts, dim3(nblk(nbox, 256)), dim3(256), 0, s, lhbx1, nbox, npbx);
    PPMG_HIP(hipGetLastError());
    return 0;
}

int rank_particles(ppmg_plan* plan, const RankGeom& g, int prec, int64_t np, const void* xp,
                   int* lpdx, int* lhbx, int64_t* icount, hipStream_t s) {
    int64_t ncells = (int64_t)g.ntot[0] * g.ntot[1] * g.ntot[2];
    // the plan's scratch is reused below: a slot table's counts are
    // overwritten, so its heads are built first
    if (materialize_heads(plan, s)) return PPMG_ERROR_FATAL;
    plan->prebinned = false;
    if (plan->key.ensure(sizeof(int) * (size_t)(np > 0 ? np : 1)) ||
        plan->slot.ensure(sizeof(int) * (size_t)(np > 0 ? np : 1)) ||
        plan->count.ensure(sizeof(int) * (size_t)ncells) || plan->err.ensure(sizeof(int) * 4))
        return PPMG_ERROR_FATAL;
    if (fill_zero(plan->count.p, sizeof(int) * (size_t)ncells, s) || fill_zero(plan->err.p, sizeof(int) * 4, s))
        return PPMG_ERROR_FATAL;
    if (np > 0) {
        unsigned gr = nblk(np, 256);
        int* st = plan->err.as<int>();
        if (g.dim == 3) {
            if (prec == PPMG_PREC_F32)
                hipLaunchKernelGGL((k_rank_keys<3, float>), dim3(gr), dim3(256), 0, s, (const float*)xp, np, g,
                                   plan->key.as<int>(), plan->slot.as<int>(), plan->count.as<int>(), st);
            else
                hipLaunchKernelGGL((k_rank_keys<3, double>), dim3(gr), dim3(256), 0, s, (const double*)xp, np, g,
                                   plan->key.as<int>(), plan->slot.as<int>(), plan->count.as<int>(), st);
        } else {
            if (prec == PPMG_PREC_F32)
                hipLaunchKernelGGL((k_rank_keys<2, float>), dim3(gr), dim3(256), 0, s, (const float*)xp, np, g,
                                   plan->key.as<int>(), plan->slot.as<int>(), plan->count.as<int>(), st);
            else
                hipLaunchKernelGGL((k_rank_keys<2, double>), dim3(gr), dim3(256), 0, s, (const double*)xp, np, g,
                                   plan->key.as<int>(), plan->slot.as<int>(), plan->count.as<int>(), st);
        }
        PPMG_HIP(hipGetLastError());
    }
    int stats[2] = {0, 0};
    PPMG_HIP(hipMemcpyAsync(stats, plan->err.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    PPMG_HIP(hipStreamSynchronize(s));
    int64_t inside = np - stats[0];
    if (exclusive_scan(plan->count.as<int>(), ncells, lhbx, plan->scan_tmp, inside, s)) return PPMG_ERROR_FATAL;
    if (np > 0) {
        hipLaunchKernelGGL(k_scatter<int>, dim3(nblk(np, 1024)), dim3(256), 0, s, plan->key.as<int>(),
                           plan->slot.as<int>(), lhbx, np, lpdx);
        hipLaunchKernelGGL(k_stabilize, dim3(nblk(ncells, 256)), dim3(256), 0, s, lhbx, ncells, lpdx);
        hipLaunchKernelGGL(k_to_one_based, dim3(nblk(inside, 256)), dim3(256), 0, s, lpdx, inside);
    }
    hipLaunchKernelGGL(k_to_one_based, dim3(nblk(ncells + 1, 256)), dim3(256), 0, s, lhbx, ncells + 1);
    PPMG_HIP(hipGetLastError());
    // the binning's keys were overwritten: recomputed from its positions when read
    plan->keys_written = false;
    *icount = inside;
    return 0;
}

}  // namespace ppmg
