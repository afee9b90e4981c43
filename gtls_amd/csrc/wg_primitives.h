// Workgroup and wavefront primitives for 64-wide CDNA4 wavefronts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gtls {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Values every lane holds identically (read back from LDS after a barrier) moved into
// scalar registers, so that branches on them compile to uniform scalar branches. Loop
// control of the persistent kernels must be uniform: a branch the compiler believes to
// be divergent can be restructured so that lanes of one wavefront reach different
// instances of a barrier.
__device__ __forceinline__ int32_t uniform(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uniform(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uniform(double x) {
    return __builtin_bit_cast(double, uniform(__builtin_bit_cast(int64_t, x)));
}
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }

// Inclusive prefix sum across the 64 lanes of a wavefront.
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        T y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}

// Exclusive prefix sum of one value per thread over the workgroup; `total` receives the
// workgroup sum. `tmp` is LDS scratch of at least blockDim.x / 64 elements.
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T x, T* tmp, T& total) {
    const int lane = lane_id(), wave = wave_id(), nwaves = blockDim.x / kWave;
    T incl = wave_inclusive_scan(x);
    if (lane == kWave - 1) tmp[wave] = incl;
    __syncthreads();
    T before = 0, sum = 0;
    for (int w = 0; w < nwaves; ++w) {
        T v = tmp[w];
        if (w < wave) before += v;
        sum += v;
    }
    total = sum;
    __syncthreads();
    return before + incl - x;
}

// Lexicographic (value, k, s) minimum: the first minimum in (k, s) order wins ties.
// Orders the LDS accesses of one wavefront before and after this point.
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct ArgMin {
    double v;
    int32_t k;
    int32_t s;
};

__device__ __forceinline__ bool argmin_less(const ArgMin& a, const ArgMin& b) {
    if (a.v < b.v) return true;
    if (b.v < a.v) return false;
    if (a.v != b.v) {            // a or b is NaN: prefer the non-NaN one
        return b.v != b.v && a.v == a.v;
    }
    if (a.k != b.k) return a.k < b.k;
    return a.s < b.s;
}

__device__ __forceinline__ ArgMin wave_argmin(ArgMin a) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        ArgMin b;
        b.v = __shfl_xor(a.v, off, kWave);
        b.k = __shfl_xor(a.k, off, kWave);
        b.s = __shfl_xor(a.s, off, kWave);
        if (argmin_less(b, a)) a = b;
    }
    return a;
}

}  // namespace gtls
