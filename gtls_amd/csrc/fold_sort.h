// Phase fold and stable phase sort of one light curve for one trial period, executed by
// one workgroup.
//
// Fold: phase = x - trunc(x), x = t / P in IEEE float64 (the fold of
// src/gputls/GPUFun.cu:28-39); the quotient is correctly rounded, the subtraction exact,
// so the keys equal the reference's bit for bit.
//
// Sort: the order of (phase, index) pairs -- the permutation a stable argsort of the
// phases gives (src/gputls/core.py:775-777). One most-significant-digit radix pass
// distributes the samples over B phase buckets (bucket = floor(phase * B), monotone in
// the phase): an LDS histogram, an LDS exclusive scan, and a scatter whose slot inside a
// bucket comes from an LDS atomic. The exact order inside each bucket is then fixed by an
// insertion sort on the full (phase, index) key, so the scatter order does not matter.
// With B ~ N/4 a bucket holds a handful of samples, so a period costs one read of t, one
// scatter and a few compares per sample instead of the 6-8 passes of a 64-bit LSD sort.
// Phases that crowd into one bucket (periods close to a multiple of the cadence) only
// make that bucket's insertion sort longer; the result is exact in every case.
#pragma once
#include "wg_primitives.h"

namespace gtls {

constexpr int kMaxBuckets = 8192;

__host__ __device__ inline int buckets_for(int64_t n) {
    int b = 256;
    while (b < kMaxBuckets && (int64_t)b * 4 < n) b <<= 1;
    return b;
}

__device__ __forceinline__ double fold_phase(double t, double P) {
    const double x = t / P;
    return x - trunc(x);
}

// Bucket of a phase in [0, 1), or of (-1, 1) when `neg` (times before zero). The map is
// monotone non-decreasing in the phase and clamped, so NaN or a wrong `neg` flag cannot
// index outside the histogram.
__device__ __forceinline__ int phase_bucket(double ph, int B, bool neg) {
    double x = neg ? (ph + 1.0) * (0.5 * (double)B) : ph * (double)B;
    x = fmin(fmax(x, 0.0), (double)(B - 1));
    return (int)x;
}

__device__ __forceinline__ bool key_greater(double ka, int32_t ia, double kb, int32_t ib) {
    return ka > kb || (ka == kb && ia > ib);
}

// Compare-exchange of two (phase, index) keys held in registers.
__device__ __forceinline__ void cmp_swap(double& ka, int32_t& ia, double& kb, int32_t& ib) {
    const bool sw = key_greater(ka, ia, kb, ib);
    const double k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const int32_t i0 = sw ? ib : ia, i1 = sw ? ia : ib;
    ka = k0; kb = k1; ia = i0; ib = i1;
}

// Batcher's odd-even merge network for 8 keys (19 compare-exchanges); unused slots hold
// +inf keys, which sort last.
__device__ __forceinline__ void sort8(double* k, int32_t* id) {
#define GTLS_CS(a, b) cmp_swap(k[a], id[a], k[b], id[b])
    GTLS_CS(0, 1); GTLS_CS(2, 3); GTLS_CS(4, 5); GTLS_CS(6, 7);
    GTLS_CS(0, 2); GTLS_CS(1, 3); GTLS_CS(4, 6); GTLS_CS(5, 7);
    GTLS_CS(1, 2); GTLS_CS(5, 6);
    GTLS_CS(0, 4); GTLS_CS(1, 5); GTLS_CS(2, 6); GTLS_CS(3, 7);
    GTLS_CS(2, 4); GTLS_CS(3, 5);
    GTLS_CS(1, 2); GTLS_CS(3, 4); GTLS_CS(5, 6);
#undef GTLS_CS
}

// Batcher's odd-even merge network for 16 keys (63 compare-exchanges).
__device__ __forceinline__ void sort16(double* k, int32_t* id) {
#define GTLS_CS(a, b) cmp_swap(k[a], id[a], k[b], id[b])
    GTLS_CS(0, 1); GTLS_CS(2, 3); GTLS_CS(0, 2); GTLS_CS(1, 3); GTLS_CS(1, 2);
    GTLS_CS(4, 5); GTLS_CS(6, 7); GTLS_CS(4, 6); GTLS_CS(5, 7); GTLS_CS(5, 6);
    GTLS_CS(0, 4); GTLS_CS(2, 6); GTLS_CS(2, 4); GTLS_CS(1, 5); GTLS_CS(3, 7);
    GTLS_CS(3, 5); GTLS_CS(1, 2); GTLS_CS(3, 4); GTLS_CS(5, 6); GTLS_CS(8, 9);
    GTLS_CS(10, 11); GTLS_CS(8, 10); GTLS_CS(9, 11); GTLS_CS(9, 10); GTLS_CS(12, 13);
    GTLS_CS(14, 15); GTLS_CS(12, 14); GTLS_CS(13, 15); GTLS_CS(13, 14); GTLS_CS(8, 12);
    GTLS_CS(10, 14); GTLS_CS(10, 12); GTLS_CS(9, 13); GTLS_CS(11, 15); GTLS_CS(11, 13);
    GTLS_CS(9, 10); GTLS_CS(11, 12); GTLS_CS(13, 14); GTLS_CS(0, 8); GTLS_CS(4, 12);
    GTLS_CS(4, 8); GTLS_CS(2, 10); GTLS_CS(6, 14); GTLS_CS(6, 10); GTLS_CS(2, 4);
    GTLS_CS(6, 8); GTLS_CS(10, 12); GTLS_CS(1, 9); GTLS_CS(5, 13); GTLS_CS(5, 9);
    GTLS_CS(3, 11); GTLS_CS(7, 15); GTLS_CS(7, 11); GTLS_CS(3, 5); GTLS_CS(7, 9);
    GTLS_CS(11, 13); GTLS_CS(1, 2); GTLS_CS(3, 4); GTLS_CS(5, 6); GTLS_CS(7, 8);
    GTLS_CS(9, 10); GTLS_CS(11, 12); GTLS_CS(13, 14);
#undef GTLS_CS
}

// Up to M keys of one bucket: load (independent loads), sort in registers, store.
template <int M>
__device__ __forceinline__ void sort_bucket_regs(double* __restrict__ key, int32_t* __restrict__ perm,
                                                 int lo, int cnt) {
    double k[M];
    int32_t id[M];
#pragma unroll
    for (int e = 0; e < M; ++e) {
        const bool in = e < cnt;
        k[e] = in ? key[lo + (in ? e : 0)] : __builtin_inf();
        id[e] = in ? perm[lo + (in ? e : 0)] : 0x7fffffff;
    }
    if constexpr (M == 8) sort8(k, id);
    else sort16(k, id);
#pragma unroll
    for (int e = 0; e < M; ++e) {
        if (e < cnt) {
            key[lo + e] = k[e];
            perm[lo + e] = id[e];
        }
    }
}

// Sorts the n samples of t by (phase, index) for period P. Writes the permutation to
// perm[0..n) and the sorted phases to key_out[0..n); key_in is n doubles of scratch.
// hist is LDS of kMaxBuckets ints, red is LDS of blockDim/64 ints. All threads call it.
__device__ void wg_fold_sort(const double* __restrict__ t, int64_t n, double P, bool neg,
                             double* __restrict__ key_in, double* __restrict__ key_out,
                             int32_t* __restrict__ perm, int* hist, int* red) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int B = buckets_for(n);
    for (int b = tid; b < B; b += nt) hist[b] = 0;
    __syncthreads();
    for (int64_t j = tid; j < n; j += nt) {
        const double ph = fold_phase(t[j], P);
        key_in[j] = ph;
        atomicAdd(&hist[phase_bucket(ph, B, neg)], 1);
    }
    __syncthreads();
    // exclusive scan of the histogram: thread-contiguous chunks, then a workgroup scan
    {
        const int per = B / nt;  // B is a multiple of 256 >= blockDim
        const int base = tid * per;
        int sum = 0;
        for (int e = 0; e < per; ++e) sum += hist[base + e];
        int total;
        int run = block_exclusive_scan(sum, red, total);
        for (int e = 0; e < per; ++e) {
            const int v = hist[base + e];
            hist[base + e] = run;
            run += v;
        }
    }
    __syncthreads();
    for (int64_t j = tid; j < n; j += nt) {
        const double ph = key_in[j];
        const int pos = atomicAdd(&hist[phase_bucket(ph, B, neg)], 1);
        perm[pos] = (int32_t)j;
        key_out[pos] = ph;
    }
    __syncthreads();
    // hist[b] now holds the end of bucket b; order each bucket by (phase, index): up to 16
    // keys in registers through a sorting network (independent loads, no dependent
    // memory round trips), larger buckets by insertion in place
    for (int b = tid; b < B; b += nt) {
        const int lo = b ? hist[b - 1] : 0;
        const int hi = hist[b];
        const int cnt = hi - lo;
        if (cnt <= 1) continue;
        if (cnt <= 8) {
            sort_bucket_regs<8>(key_out, perm, lo, cnt);
            continue;
        }
        if (cnt <= 16) {
            sort_bucket_regs<16>(key_out, perm, lo, cnt);
            continue;
        }
        for (int i = lo + 1; i < hi; ++i) {
            const double k = key_out[i];
            const int32_t id = perm[i];
            int j = i - 1;
            while (j >= lo && key_greater(key_out[j], perm[j], k, id)) {
                key_out[j + 1] = key_out[j];
                perm[j + 1] = perm[j];
                --j;
            }
            key_out[j + 1] = k;
            perm[j + 1] = id;
        }
    }
    __syncthreads();
}

}  // namespace gtls
