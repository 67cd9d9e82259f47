// Stateless Philox4x32-10 generators (counter = (index, call, 0, 0), key = seed) and fills.
// Reference: include/utils/device_fields.cuh:150-298.  The uniform draws are bit-exact
// with the published Random123 algorithm; the normal draws use Box-Muller with the
// device's log/cos/sin.
#include "internal.hpp"

namespace dpmd {
namespace {

__device__ __forceinline__ void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double u01d(uint32_t hi, uint32_t lo) {
    const uint64_t bits = ((uint64_t)hi << 21) | ((uint64_t)lo >> 11);
    return ((double)bits + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ void k_uniform_2d(double* ox, double* oy, int n, uint64_t seed, uint64_t call, double lox, double hix,
                             double loy, double hiy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c0 = (uint32_t)i, c1 = (uint32_t)call, c2 = 0u, c3 = 0u;
    philox10(c0, c1, c2, c3, (uint32_t)(seed & 0xFFFFFFFFull), (uint32_t)(seed >> 32));
    double ux = u01d(c0, c1); if (ux == 1.0) ux = 0.0;
    double uy = u01d(c2, c3); if (uy == 1.0) uy = 0.0;
    ox[i] = lox + (hix - lox) * ux;
    oy[i] = loy + (hiy - loy) * uy;
}

__global__ void k_uniform_1d(double* o, int n, uint64_t seed, uint64_t call, double lo, double hi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c0 = (uint32_t)i, c1 = (uint32_t)call, c2 = 0u, c3 = 0u;
    philox10(c0, c1, c2, c3, (uint32_t)(seed & 0xFFFFFFFFull), (uint32_t)(seed >> 32));
    double u = u01d(c0, c1); if (u == 1.0) u = 0.0;
    o[i] = lo + (hi - lo) * u;
}

__global__ void k_normal_2d(double* ox, double* oy, int n, uint64_t seed, uint64_t call, double mx, double sx,
                            double my, double sy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c0 = (uint32_t)i, c1 = (uint32_t)call, c2 = 0u, c3 = 0u;
    philox10(c0, c1, c2, c3, (uint32_t)(seed & 0xFFFFFFFFull), (uint32_t)(seed >> 32));
    const double u0 = u01d(c0, c1), u1 = u01d(c2, c3);
    const double r = sqrt(-2.0 * log(u0));
    const double th = 6.28318530717958647692 * u1;
    ox[i] = mx + sx * (r * cos(th));
    oy[i] = my + sy * (r * sin(th));
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_f64(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

void copy_f64(hipStream_t st, double* dst, const double* src, int64_t n) {
    if (n <= 0) return;
    k_copy_f64<<<(unsigned)((n + DPMD_BLOCK - 1) / DPMD_BLOCK), DPMD_BLOCK, 0, st>>>(dst, src, n);
    HIP_OK(hipGetLastError());
}

void rand_uniform_2d(hipStream_t st, double* ox, double* oy, int n, uint64_t seed, uint64_t call, double lox,
                     double hix, double loy, double hiy) {
    if (n <= 0) return;
    k_uniform_2d<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(ox, oy, n, seed, call, lox, hix, loy, hiy);
    HIP_OK(hipGetLastError());
}
void rand_uniform_1d(hipStream_t st, double* o, int n, uint64_t seed, uint64_t call, double lo, double hi) {
    if (n <= 0) return;
    k_uniform_1d<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(o, n, seed, call, lo, hi);
    HIP_OK(hipGetLastError());
}
void rand_normal_2d(hipStream_t st, double* ox, double* oy, int n, uint64_t seed, uint64_t call, double mx, double sx,
                    double my, double sy) {
    if (n <= 0) return;
    k_normal_2d<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(ox, oy, n, seed, call, mx, sx, my, sy);
    HIP_OK(hipGetLastError());
}
void fill_f64(hipStream_t st, double* p, int64_t n, double v) {
    if (n <= 0) return;
    k_fill_f64<<<(unsigned)((n + DPMD_BLOCK - 1) / DPMD_BLOCK), DPMD_BLOCK, 0, st>>>(p, n, v);
    HIP_OK(hipGetLastError());
}
void fill_i32(hipStream_t st, int* p, int64_t n, int v) {
    if (n <= 0) return;
    k_fill_i32<<<(unsigned)((n + DPMD_BLOCK - 1) / DPMD_BLOCK), DPMD_BLOCK, 0, st>>>(p, n, v);
    HIP_OK(hipGetLastError());
}

}  // namespace dpmd
