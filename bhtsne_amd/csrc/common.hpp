// Shared device/host helpers for the bhtsne MI355X engine (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace bht {

// ---- error handling ------------------------------------------------------------------------

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
    OK = 0,
    ERR_INVALID = 1,  // invalid argument: same message text as the reference's panic
    ERR_HIP = 2,      // HIP runtime failure
    ERR_OOM = 3,      // allocation failed / HBM budget exceeded
    ERR_STATE = 4,    // call out of order (e.g. run before data)
};

#define BHT_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::bht::Error(e_ == hipErrorOutOfMemory ? ::bht::ERR_OOM : ::bht::ERR_HIP,     \
                               std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

#define BHT_LAUNCH_CHECK() BHT_HIP(hipGetLastError())

constexpr int WAVE = 64;

inline unsigned ceil_div(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

// ---- device helpers ------------------------------------------------------------------------

// Pointers read from a device table are generic, and the compiler turns accesses through them
// into flat instructions. A flat load counts on the LDS counter as well, so every wait for an LDS
// read would also wait for all outstanding flat loads. Kernels that mix LDS reads with global
// loads access global memory through these address-space-qualified views instead.
#define BHT_GLOBAL __attribute__((address_space(1)))
#define BHT_LDS __attribute__((address_space(3)))
template <class T>
__device__ __forceinline__ BHT_GLOBAL T *global_view(T *p) {
    return (BHT_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ BHT_LDS T *lds_view(T *p) {
    return (BHT_LDS T *)p;
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-wide sum of a double held by every thread; result valid in thread 0.  `scratch` needs
// blockDim/64 entries.
__device__ __forceinline__ double block_sum_f64(double v, double *scratch) {
    v = wave_sum_f64(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += scratch[i];
    __syncthreads();
    return r;
}

// "Last block" detection for single-pass reductions: every block publishes its partial, then
// increments `counter`; the block that sees gridDim-1 combines all partials in block order (so the
// result does not depend on scheduling) and resets the counter for the next launch.
__device__ __forceinline__ bool arrive_last(unsigned *counter, unsigned nblocks) {
    __shared__ bool last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned prev = atomicAdd(counter, 1u);
        last = (prev == nblocks - 1);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// Rust f32::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives.
__device__ __forceinline__ float rust_signum(float x) {
    return isnan(x) ? x : copysignf(1.0f, x);
}

}  // namespace bht
