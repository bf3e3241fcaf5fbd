// ppmg_gen.hip -- device generators of the synthetic workloads, bit-compatible
// with ppm_legacy_amd/workloads.py for the uniforms and lattice positions
// (Philox4x32-10, Salmon et al. SC'11; 53-bit uniforms from two words).
#include "ppmg_internal.h"

namespace ppmg {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        uint64_t p0 = (uint64_t)M0 * c[0];
        uint64_t p1 = (uint64_t)M1 * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

__device__ __forceinline__ double res53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

__device__ __forceinline__ void uniforms2(uint64_t seed, uint32_t stream, uint64_t i, double& ua, double& ub) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), stream, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    ua = res53(c[0], c[1]);
    ub = res53(c[2], c[3]);
}

// workloads.jittered_lattice: x = xmin + ((cell + 0.5) + 2 jitter (u - 0.5)) h
__global__ void __launch_bounds__(256) k_gen_lattice(int dim, int N, double jitter, uint64_t seed, double* xp,
                                                     int64_t n) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[4];
    uniforms2(seed, 0u, (uint64_t)i, u[0], u[1]);
    if (dim == 3) uniforms2(seed, 1u, (uint64_t)i, u[2], u[3]);
    int64_t cell[3] = {i % N, (i / N) % N, i / ((int64_t)N * N)};
    double h = 1.0 / (double)N;
    for (int d = 0; d < dim; ++d) {
        double t = (2.0 * jitter) * (u[d] - 0.5);
        xp[i * dim + d] = 0.0 + (((double)cell[d] + 0.5) + t) * h;
    }
}

// workloads.normals (Box-Muller cosine branch); stream 16 + component.
template <typename T>
__global__ void __launch_bounds__(256) k_gen_normals(int64_t count, int lda, uint64_t seed, T* out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * lda) return;
    int64_t i = t / lda;
    int c = (int)(t % lda);
    double ua, ub;
    uniforms2(seed, 16u + (uint32_t)c, (uint64_t)i, ua, ub);
    double z = sqrt(-2.0 * log1p(-ua)) * cos(2.0 * M_PI * ub);
    out[t] = (T)z;
}

// workloads.divfree_velocity: u(x) = sum_m (A_m x k_m) cos(2 pi k_m.x + phi_m)
__global__ void __launch_bounds__(256) k_gen_divfree(int64_t np, const double* __restrict__ xp, int nmodes,
                                                     const double* __restrict__ k, const double* __restrict__ A,
                                                     const double* __restrict__ phi, double* __restrict__ u) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double x = xp[p * 3], y = xp[p * 3 + 1], z = xp[p * 3 + 2];
    double ux = 0, uy = 0, uz = 0;
    for (int m = 0; m < nmodes; ++m) {
        double kx = k[3 * m], ky = k[3 * m + 1], kz = k[3 * m + 2];
        double ax = A[3 * m], ay = A[3 * m + 1], az = A[3 * m + 2];
        double bx = ay * kz - az * ky, by = az * kx - ax * kz, bz = ax * ky - ay * kx;
        double c = cos(2.0 * M_PI * (x * kx + y * ky + z * kz) + phi[m]);
        ux += c * bx;
        uy += c * by;
        uz += c * bz;
    }
    u[p * 3] = ux;
    u[p * 3 + 1] = uy;
    u[p * 3 + 2] = uz;
}

}  // namespace ppmg

using namespace ppmg;

extern "C" {

int ppmg_gen_jittered_lattice(int dim, int n, double jitter, uint64_t seed, double* xp, void* stream) {
    int64_t np = 1;
    for (int d = 0; d < dim; ++d) np *= n;
    unsigned g = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_gen_lattice, dim3(g), dim3(256), 0, (hipStream_t)stream, dim, n, jitter, seed, xp, np);
    PPMG_HIP(hipGetLastError());
    return 0;
}

int ppmg_gen_normals(int64_t count, int lda, uint64_t seed, int prec, void* out, void* stream) {
    int64_t tot = count * lda;
    unsigned g = (unsigned)((tot + 255) / 256);
    if (prec == PPMG_PREC_F64)
        hipLaunchKernelGGL((k_gen_normals<double>), dim3(g), dim3(256), 0, (hipStream_t)stream, count, lda, seed,
                           (double*)out);
    else
        hipLaunchKernelGGL((k_gen_normals<float>), dim3(g), dim3(256), 0, (hipStream_t)stream, count, lda, seed,
                           (float*)out);
    PPMG_HIP(hipGetLastError());
    return 0;
}

int ppmg_gen_divfree_velocity(int64_t np, const double* xp, int nmodes, const double* k, const double* A,
                              const double* phi, double* u, void* stream) {
    unsigned g = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_gen_divfree, dim3(g), dim3(256), 0, (hipStream_t)stream, np, xp, nmodes, k, A, phi, u);
    PPMG_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
