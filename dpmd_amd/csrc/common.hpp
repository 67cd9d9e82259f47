// Shared definitions for the dpmd_amd HIP library (gfx950, FP64, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>

#define DPMD_BLOCK 256
// largest system the persistent per-system FIRE kernel takes (64 chunks of 256)
constexpr int kPersistMaxN = 16384;

namespace dpmd {

struct Error : std::runtime_error { using std::runtime_error::runtime_error; };
struct ArgError : Error { using Error::Error; };        // maps to DPMD_ERR_ARGUMENT
struct StateError : Error { using Error::Error; };      // maps to DPMD_ERR_STATE
struct OverflowError : Error { using Error::Error; };   // maps to DPMD_ERR_OVERFLOW

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw ::dpmd::Error(std::string(#expr) + ": " + hipGetErrorString(_e));           \
    } while (0)

// Owning device vector (not copyable).  Grows only; never shrinks while alive.
template <class T>
struct DVec {
    T* p = nullptr;
    int64_t n = 0;
    int64_t cap = 0;
    DVec() = default;
    DVec(const DVec&) = delete;
    DVec& operator=(const DVec&) = delete;
    ~DVec() { release(); }
    void resize(int64_t m) {
        if (m > cap) {
            if (p) HIP_OK(hipFree(p));
            p = nullptr;
            const int64_t bytes = (m > 0 ? m : 1) * (int64_t)sizeof(T);
            HIP_OK(hipMalloc((void**)&p, bytes));
            cap = m;
        }
        n = m;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0; cap = 0;
    }
    void zero(hipStream_t st) { if (n) HIP_OK(hipMemsetAsync(p, 0, n * sizeof(T), st)); }
    void upload(const T* h, int64_t m, hipStream_t st) {
        resize(m);
        if (m) HIP_OK(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, st));
        HIP_OK(hipStreamSynchronize(st));   // the host buffer may be freed on return
    }
    void download(T* h, hipStream_t st) const {
        if (n) HIP_OK(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
    }
};

inline int blocks_for(int64_t n, int b = DPMD_BLOCK) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------------------------------
// Deterministic reductions.  A wave of 64 lanes folds with an xor butterfly (offsets
// 32..1, v <- v + v[lane^off]); a 256-thread block folds its 4 wave sums as
// ((w0+w1)+(w2+w3)).  Floating-point addition is commutative, so every lane of a pair
// holds the same bits after each butterfly stage, and the result does not depend on
// scheduling.  The host restatement of the test suite folds in exactly this order.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
// Every thread of the 256-thread block must call; every thread gets the result.
__device__ __forceinline__ double block_sum256(double v, double* lds4) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds4[w] = v;
    __syncthreads();
    const double r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}
__device__ __forceinline__ double block_max256(double v, double* lds4) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds4[w] = v;
    __syncthreads();
    const double r = fmax(fmax(lds4[0], lds4[1]), fmax(lds4[2], lds4[3]));
    __syncthreads();
    return r;
}

// Workgroups are handed to the 8 XCDs round-robin (block b -> XCD b%8).  Remap so that
// each XCD works on one contiguous range of chunks: spatially sorted particles then keep
// their neighbor gathers inside that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, slot = b >> 3;
    return x * q + (x < r ? x : r) + slot;
}

// ---------------------------------------------------------------------------------------
// Geometry (same expressions as include/kernels/base_particle_kernels.cuh:117-179)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wrap01(double u) {
    u -= floor(u);
    return (u >= 1.0) ? 0.0 : u;
}
__device__ __forceinline__ double min_image(double d, double L, double Linv) {
    return (L > 0.0) ? (d - L * rint(d * Linv)) : d;
}

// sin and cos of a rotation increment (reference: sincos in rigid_bumpy.cu:51).  Cody-Waite
// reduction by pi/2 with the fdlibm two-part constant, Taylor series to x^17 / x^18 on
// |r| <= pi/4 in Horner form; within 1 ulp of libm on the tested ranges.  Written out
// (no library call) so host and device produce the same bits.
__device__ __forceinline__ void own_sincos(double x, double* sp, double* cp) {
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    const double two_over_pi = 6.36619772367581382433e-01;
    const double k = rint(x * two_over_pi);
    const double r = (x - k * pio2_1) - k * pio2_1t;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;
    ps = -1.0 / 1307674368000.0 + r2 * ps;
    ps = 1.0 / 6227020800.0 + r2 * ps;
    ps = -1.0 / 39916800.0 + r2 * ps;
    ps = 1.0 / 362880.0 + r2 * ps;
    ps = -1.0 / 5040.0 + r2 * ps;
    ps = 1.0 / 120.0 + r2 * ps;
    ps = -1.0 / 6.0 + r2 * ps;
    const double s = r + (r * r2) * ps;
    double pc = -1.0 / 6402373705728000.0;
    pc = 1.0 / 20922789888000.0 + r2 * pc;
    pc = -1.0 / 87178291200.0 + r2 * pc;
    pc = 1.0 / 479001600.0 + r2 * pc;
    pc = -1.0 / 3628800.0 + r2 * pc;
    pc = 1.0 / 40320.0 + r2 * pc;
    pc = -1.0 / 720.0 + r2 * pc;
    pc = 1.0 / 24.0 + r2 * pc;
    const double c = (1.0 - 0.5 * r2) + (r2 * r2) * pc;
    const long long q = ((long long)k) & 3;
    if (q == 0) { *sp = s; *cp = c; }
    else if (q == 1) { *sp = c; *cp = -s; }
    else if (q == 2) { *sp = -s; *cp = -c; }
    else { *sp = -c; *cp = s; }
}

}  // namespace dpmd
