// Phase fold and stable phase sort of one light curve for one trial period, executed by
// one workgroup.
//
// Fold: phase = x - trunc(x), x = t / P in IEEE float64 (the fold of
// src/gputls/GPUFun.cu:28-39); the quotient is correctly rounded, the subtraction exact,
// so the keys equal the reference's bit for bit. The windowed form forms the quotient from
// the period's reciprocal (fold_math.h: the same phases); a curve with a time below 2^-960
// in magnitude, or a period outside [2^-1000, 2^1000], takes the global form, which
// divides.
//
// Sort: the order of (phase, index) pairs -- the permutation a stable argsort of the
// phases gives (src/gputls/core.py:775-777). One most-significant-digit radix pass
// distributes the samples over B phase buckets (bucket = floor(phase * B), monotone in
// the phase): an LDS histogram, an LDS exclusive scan, and a scatter whose slot inside a
// bucket comes from an LDS atomic. The exact order inside each bucket is then fixed on the
// full (phase, index) key, so the scatter order does not matter. With B ~ N/4 a bucket
// holds a handful of samples (a register sorting network per thread); phases that crowd
// into one bucket (periods close to a multiple of the cadence: tens to hundreds of
// samples) are sorted by one wavefront with a bitonic network in registers, larger ones
// by the workgroup in LDS, and the largest (e.g. repeated times) by LDS-sorted chunks and
// rank-based merges. The cost of a bucket is bounded by m log^2 m in every case.
//
// Windowed form (wg_fold_sort_windows, the usual path): after the histogram, the buckets
// are cut into windows of about kWinTarget samples and one pass sends every sample index
// to its window's range (a few dozen ranges, so partly written lines stay in cache until
// they fill); for each window the workgroup folds its samples again, stages them in LDS at
// their bucket positions, orders every bucket there and hands the window to a sink. A
// period whose largest bucket exceeds kWaveSortMax keys takes the global form above.
#pragma once
#include "fold_math.h"
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

// The same phase with the quotient from rP = fold_reciprocal(P), for times that are zero
// or at least 2^-960 in magnitude.
__device__ __forceinline__ double fold_phase_r(double t, double P, double rP) {
    const double x = fold_quotient(t, P, rP);
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

// Compare-exchange of two (phase, index) keys held in registers; with PhaseOnly, on the
// phase alone (equal phases may then end in either order), the phases placed by min/max.
template <bool PhaseOnly = false>
__device__ __forceinline__ void cmp_swap(double& ka, int32_t& ia, double& kb, int32_t& ib) {
    const bool sw = PhaseOnly ? ka > kb : key_greater(ka, ia, kb, ib);
    const double k0 = PhaseOnly ? fmin(ka, kb) : (sw ? kb : ka);
    const double k1 = PhaseOnly ? fmax(ka, kb) : (sw ? ka : kb);
    const int32_t i0 = sw ? ib : ia, i1 = sw ? ia : ib;
    ka = k0; kb = k1; ia = i0; ib = i1;
}

// Batcher's odd-even merge network for 8 keys (19 compare-exchanges); unused slots hold
// +inf keys, which sort last.
template <bool PhaseOnly = false>
__device__ __forceinline__ void sort8(double* k, int32_t* id) {
#define GTLS_CS(a, b) cmp_swap<PhaseOnly>(k[a], id[a], k[b], id[b])
    GTLS_CS(0, 1); GTLS_CS(2, 3); GTLS_CS(4, 5); GTLS_CS(6, 7);
    GTLS_CS(0, 2); GTLS_CS(1, 3); GTLS_CS(4, 6); GTLS_CS(5, 7);
    GTLS_CS(1, 2); GTLS_CS(5, 6);
    GTLS_CS(0, 4); GTLS_CS(1, 5); GTLS_CS(2, 6); GTLS_CS(3, 7);
    GTLS_CS(2, 4); GTLS_CS(3, 5);
    GTLS_CS(1, 2); GTLS_CS(3, 4); GTLS_CS(5, 6);
#undef GTLS_CS
}

// Batcher's odd-even merge network for 16 keys (63 compare-exchanges); with N < 16 the
// network for the first N inputs, the others +inf: comparators that touch an input >= N
// are dropped (the +inf keys start at the top and only ever move up), 42 for N = 12.
template <bool PhaseOnly = false, int N = 16>
__device__ __forceinline__ void sort16(double* k, int32_t* id) {
#define GTLS_CS(a, b) if constexpr ((b) < N) cmp_swap<PhaseOnly>(k[a], id[a], k[b], id[b])
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
    // on the phases alone, then again on the full key only when two of the bucket's
    // phases are equal (their order is then decided by the index)
    if constexpr (M == 8) sort8<true>(k, id);
    else sort16<true, M>(k, id);
    bool tie = false;
#pragma unroll
    for (int e = 0; e + 1 < M; ++e) tie |= (e + 1 < cnt) && k[e] == k[e + 1];
    if (tie) {
        if constexpr (M == 8) sort8(k, id);
        else sort16<false, M>(k, id);
    }
#pragma unroll
    for (int e = 0; e < M; ++e) {
        if (e < cnt) {
            key[lo + e] = k[e];
            perm[lo + e] = id[e];
        }
    }
}

constexpr int kWaveSortMax = 256;    // buckets up to this size: one wavefront, registers
constexpr int kBlockSortMax = 2048;  // up to this size: the workgroup, in LDS

// Bitonic sort of cnt <= 64 E keys by one wavefront: element i = lane E + e sits in
// register e of its lane; partners i ^ j with j < E are in the same lane, the others one
// shuffle away. Padding keys (+inf, INT_MAX) sort last.
template <int E>
__device__ __forceinline__ void wave_bitonic(double* __restrict__ key, int32_t* __restrict__ id,
                                             int64_t lo, int cnt) {
    const int lane = lane_id();
    double k[E];
    int32_t x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = lane * E + e;
        const bool in = i < cnt;
        k[e] = in ? key[lo + (in ? i : 0)] : __builtin_inf();
        x[e] = in ? id[lo + (in ? i : 0)] : 0x7fffffff;
    }
    constexpr int M = 64 * E;
#pragma unroll
    for (int kk = 2; kk <= M; kk <<= 1) {
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            if (j >= E) {
                const int lj = j / E;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = lane * E + e;
                    const double pk = __shfl_xor(k[e], lj, kWave);
                    const int32_t px = __shfl_xor(x[e], lj, kWave);
                    const bool want_min = ((i & j) == 0) == ((i & kk) == 0);
                    const bool pg = key_greater(pk, px, k[e], x[e]);   // partner > mine
                    const bool take = want_min ? !pg : pg;              // partner's is the pick
                    if (take) {
                        k[e] = pk;
                        x[e] = px;
                    }
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if ((e & j) == 0) {
                        const int i = lane * E + e;
                        const bool asc = (i & kk) == 0;
                        if (key_greater(k[e], x[e], k[e | j], x[e | j]) == asc) {
                            const double tk = k[e];
                            const int32_t tx = x[e];
                            k[e] = k[e | j];
                            x[e] = x[e | j];
                            k[e | j] = tk;
                            x[e | j] = tx;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = lane * E + e;
        if (i < cnt) {
            key[lo + i] = k[e];
            id[lo + i] = x[e];
        }
    }
}

// Bitonic sort of cnt <= kBlockSortMax keys by the whole workgroup in LDS (sk: doubles,
// si: ints, kBlockSortMax each). All threads call it with the same arguments.
__device__ void block_bitonic(double* __restrict__ key, int32_t* __restrict__ id, int64_t lo,
                              int cnt, double* sk, int32_t* si) {
    const int tid = threadIdx.x, nt = blockDim.x;
    int M = 64;
    while (M < cnt) M <<= 1;
    for (int i = tid; i < M; i += nt) {
        const bool in = i < cnt;
        sk[i] = in ? key[lo + (in ? i : 0)] : __builtin_inf();
        si[i] = in ? id[lo + (in ? i : 0)] : 0x7fffffff;
    }
    __syncthreads();
    for (int kk = 2; kk <= M; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < M; i += nt) {
                const int p = i ^ j;
                if (p > i) {
                    const bool asc = (i & kk) == 0;
                    if (key_greater(sk[i], si[i], sk[p], si[p]) == asc) {
                        const double tk = sk[i];
                        const int32_t tx = si[i];
                        sk[i] = sk[p];
                        si[i] = si[p];
                        sk[p] = tk;
                        si[p] = tx;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < cnt; i += nt) {
        key[lo + i] = sk[i];
        id[lo + i] = si[i];
    }
    __syncthreads();
}

// Number of keys in the sorted run (rk, ri)[0, len) less than (k, x).
__device__ __forceinline__ int rank_in(const double* rk, const int32_t* ri, int len, double k,
                                       int32_t x) {
    int a = 0, b = len;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (key_greater(k, x, rk[m], ri[m])) a = m + 1; else b = m;
    }
    return a;
}

// Sort of any bucket size by the workgroup: kBlockSortMax-key chunks in LDS, then merge
// passes between (key, id) and the scratch (tk, ti) in which every key's place comes from
// its rank in the other run (keys are unique: the index breaks phase ties).
__device__ void block_merge_sort(double* __restrict__ key, int32_t* __restrict__ id,
                                 double* __restrict__ tk, int32_t* __restrict__ ti, int64_t lo,
                                 int cnt, double* sk, int32_t* si) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int c0 = 0; c0 < cnt; c0 += kBlockSortMax)
        block_bitonic(key, id, lo + c0, min(kBlockSortMax, cnt - c0), sk, si);
    double* sk_ = key + lo;
    int32_t* si_ = id + lo;
    double* dk = tk + lo;
    int32_t* di = ti + lo;
    for (int w = kBlockSortMax; w < cnt; w <<= 1) {
        for (int i = tid; i < cnt; i += nt) {
            const int a0 = (i / (2 * w)) * (2 * w);
            const int b0 = min(a0 + w, cnt), b1 = min(a0 + 2 * w, cnt);
            const double k = sk_[i];
            const int32_t x = si_[i];
            int pos;
            if (i < b0) pos = a0 + (i - a0) + rank_in(sk_ + b0, si_ + b0, b1 - b0, k, x);
            else pos = a0 + (i - b0) + rank_in(sk_ + a0, si_ + a0, b0 - a0, k, x);
            dk[pos] = k;
            di[pos] = x;
        }
        __syncthreads();
        double* t0 = sk_; sk_ = dk; dk = t0;
        int32_t* t1 = si_; si_ = di; di = t1;
    }
    if (sk_ != key + lo) {
        for (int i = tid; i < cnt; i += nt) {
            key[lo + i] = sk_[i];
            id[lo + i] = si_[i];
        }
        __syncthreads();
    }
}

// Sorts the n samples of t by (phase, index) for period P. Writes the permutation to
// perm[0..n) and the sorted phases to key_out[0..n). Scratch: tmp_key, tmp_id (n each) and
// lists (kMaxBuckets int64). LDS: hist (kMaxBuckets ints, 16-byte aligned), red
// (blockDim/64 ints), counts (2 ints). All threads call it.
__device__ __forceinline__ void wg_fold_sort(const double* __restrict__ t, int64_t n, double P, bool neg,
                             double* __restrict__ key_out, int32_t* __restrict__ perm,
                             double* __restrict__ tmp_key, int32_t* __restrict__ tmp_id,
                             int64_t* __restrict__ lists, int* hist, int* red, int* counts) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int B = buckets_for(n);
    for (int b = tid; b < B; b += nt) hist[b] = 0;
    if (tid < 2) counts[tid] = 0;
    __syncthreads();
    constexpr int U = 4;   // samples in flight per thread
    for (int64_t base = 0; base < n; base += (int64_t)U * nt) {
        double tv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = base + (int64_t)u * nt + tid;
            tv[u] = j < n ? t[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = base + (int64_t)u * nt + tid;
            if (j < n) atomicAdd(&hist[phase_bucket(fold_phase(tv[u], P), B, neg)], 1);
        }
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
    // scatter, folding again (t is re-read from cache rather than phases stored)
    for (int64_t base = 0; base < n; base += (int64_t)U * nt) {
        double tv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = base + (int64_t)u * nt + tid;
            tv[u] = j < n ? t[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = base + (int64_t)u * nt + tid;
            if (j < n) {
                const double ph = fold_phase(tv[u], P);
                const int pos = atomicAdd(&hist[phase_bucket(ph, B, neg)], 1);
                perm[pos] = (int32_t)j;
                key_out[pos] = ph;
            }
        }
    }
    __syncthreads();
    // hist[b] now holds the end of bucket b; order each bucket by (phase, index): up to 16
    // keys in registers through a sorting network, larger buckets are listed (from the
    // front up to kWaveSortMax keys, from the back beyond) for the passes below
    for (int b = tid; b < B; b += nt) {
        const int lo = b ? hist[b - 1] : 0;
        const int hi = hist[b];
        const int cnt = hi - lo;
        if (cnt <= 1) continue;
        if (cnt <= 8) {
            sort_bucket_regs<8>(key_out, perm, lo, cnt);
        } else if (cnt <= 16) {
            sort_bucket_regs<16>(key_out, perm, lo, cnt);
        } else if (cnt <= kWaveSortMax) {
            const int e = atomicAdd(&counts[0], 1);
            lists[e] = ((int64_t)lo << 32) | (int64_t)cnt;
        } else {
            const int e = atomicAdd(&counts[1], 1);
            lists[kMaxBuckets - 1 - e] = ((int64_t)lo << 32) | (int64_t)cnt;
        }
    }
    __syncthreads();
    const int nmid = counts[0], nbig = counts[1];
    // 17 .. kWaveSortMax keys: one wavefront each
    {
        const int wave = uniform((int32_t)wave_id());
        const int nwaves = nt / kWave;
        for (int e = wave; e < nmid; e += nwaves) {
            const int64_t ent = uniform(lists[e]);
            const int64_t lo = ent >> 32;
            const int cnt = (int)(ent & 0xffffffffLL);
            if (cnt <= 64) wave_bitonic<1>(key_out, perm, lo, cnt);
            else if (cnt <= 128) wave_bitonic<2>(key_out, perm, lo, cnt);
            else wave_bitonic<4>(key_out, perm, lo, cnt);
        }
    }
    __syncthreads();
    // larger buckets: the workgroup, in LDS (the histogram is no longer needed)
    double* sk = (double*)hist;
    int32_t* si = (int32_t*)(sk + kBlockSortMax);
    for (int e = 0; e < nbig; ++e) {
        const int64_t ent = uniform(lists[kMaxBuckets - 1 - e]);
        const int64_t lo = ent >> 32;
        const int cnt = (int)(ent & 0xffffffffLL);
        if (cnt <= kBlockSortMax) block_bitonic(key_out, perm, lo, cnt, sk, si);
        else block_merge_sort(key_out, perm, tmp_key, tmp_id, lo, cnt, sk, si);
    }
    __syncthreads();
}

// The windowed form keeps bucket b's count / position at hist[hslot(b)]: one spare int per
// 32 buckets, so that threads scanning consecutive 32-bucket runs start in different LDS
// banks (33 is odd). Its histogram array holds kHistLen ints.
constexpr int kHistLen = kMaxBuckets + kMaxBuckets / 32;
__device__ __forceinline__ int hslot(int b) { return b + (b >> 5); }

constexpr int kWinCap = 2816;                        // samples one window stages in LDS
constexpr int kWinTarget = kWinCap - kWaveSortMax;   // window size before its last bucket
constexpr int kWinMid = kWinCap / 17 + 1;            // buckets of 17+ keys in one window
constexpr int kWinPer = (kWinCap + 255) / 256;       // staged samples per thread (256 threads)
constexpr int kMaxWin = 255;                         // windows of one period (uint8 table)

// Windowed stable phase sort of the n samples of t for period P (see the file comment).
// Scratch: binned (n ints: sample indices grouped by window). LDS: hist (kHistLen ints), red (blockDim/64 ints), counts (4 ints), sphi /
// sidx (kWinCap doubles / ints), mids (kWinMid int64), wtab (kMaxBuckets bytes: window of
// each bucket), wb / ws (kMaxWin + 1 ints: first bucket / first position of each window),
// wcur (kMaxWin ints).
// Runs with 256 threads (kWinPer staged samples per thread).
// sink(wstart, wcnt) is called by every thread for each window: the samples of sorted
// positions wstart .. wstart + wcnt - 1 sit in sphi / sidx[0 .. wcnt); it may use
// barriers. Returns false (every thread) when a bucket holds more than kWaveSortMax keys,
// the period needs more than kMaxWin windows, or the reciprocal form of the fold does not
// apply (a time of the curve below 2^-960 in magnitude, a period outside [2^-1000,
// 2^1000]): nothing was sorted, the caller takes the global form.
template <class Sink>
__device__ __forceinline__ bool wg_fold_sort_windows(
    const double* __restrict__ t, int64_t n, double P, bool neg,
    int32_t* __restrict__ binned, int* hist, int* red, int* counts, double* sphi,
    int32_t* sidx, int64_t* mids, uint8_t* wtab, int* wb, int* ws, int* wcur, Sink sink) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int B = buckets_for(n);
    const double rP = fold_reciprocal(P);
    for (int b = tid; b < B; b += nt) hist[hslot(b)] = 0;
    if (tid < 4) counts[tid] = 0;
    __syncthreads();
    constexpr int U = 16;   // samples in flight per thread
    bool small = rP != rP;
    for (int64_t base = 0; base < n; base += (int64_t)U * nt) {
        double tv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = base + (int64_t)u * nt + tid;
            tv[u] = j < n ? t[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t j = base + (int64_t)u * nt + tid;
            if (j < n) {
                small |= fold_time_small(tv[u]);
                atomicAdd(&hist[hslot(phase_bucket(fold_phase_r(tv[u], P, rP), B, neg))], 1);
            }
        }
    }
    if (small) atomicOr(&counts[0], 2);
    __syncthreads();
    // exclusive scan of the histogram; flag a bucket too large for the windowed form
    {
        const int per = B / nt;  // B is a multiple of 256 >= blockDim
        const int base = tid * per;
        int sum = 0, mx = 0;
        for (int e = 0; e < per; ++e) {
            const int v = hist[hslot(base + e)];
            sum += v;
            mx = max(mx, v);
        }
        int total;
        int run = block_exclusive_scan(sum, red, total);
        for (int e = 0; e < per; ++e) {
            const int v = hist[hslot(base + e)];
            hist[hslot(base + e)] = run;
            run += v;
        }
        if (mx > kWaveSortMax) atomicOr(&counts[0], 1);
    }
    __syncthreads();
    // Every thread reads the flags before any thread leaves: the caller's global form resets
    // counts[] before its first barrier, so a late read after a return could see 0.
    {
        const bool too_big = counts[0] != 0;
        __syncthreads();
        if (too_big) return false;
    }
    // the windows: buckets wb[w] .. wb[w + 1] - 1, wb[w + 1] the first bucket starting at or
    // past ws[w] + kWinTarget, so that a window holds fewer than kWinCap samples. The first
    // wavefront finds each boundary by 64-way probes of the bucket starts (monotone): three
    // rounds for 8192 buckets instead of a 13-step chain of dependent reads.
    if (uniform((int32_t)wave_id()) == 0) {
        const int lane = lane_id();
        int K = 0, bfirst = 0, wstart = 0;
        while (bfirst < B && K < kMaxWin) {
            // first b in [bfirst + 1, B) with hist[b] >= target, else B; the answer stays
            // in [lo, hi] with hi == B or hist[hi] >= target
            const int target = wstart + kWinTarget;
            int lo = bfirst + 1, hi = B;
            while (lo < hi) {
                const int step = (hi - lo + kWave - 1) / kWave;
                const int b = lo + lane * step;
                const bool ge = b >= hi || hist[hslot(min(b, B - 1))] >= target;
                const unsigned long long m = __ballot(ge);
                if (m == 0ULL) {
                    lo = lo + (kWave - 1) * step + 1;   // every probe below the target
                } else {
                    const int f = __ffsll((long long)m) - 1;
                    if (f == 0) {
                        hi = lo;
                    } else {
                        hi = min(hi, lo + f * step);
                        lo = lo + (f - 1) * step + 1;
                    }
                }
            }
            if (lane == 0) {
                wb[K] = bfirst;
                ws[K] = wstart;
            }
            ++K;
            bfirst = lo;
            wstart = lo < B ? hist[hslot(lo)] : (int)n;
        }
        if (lane == 0) {
            wb[K] = B;
            ws[K] = (int)n;
            counts[2] = K;
            counts[3] = bfirst < B ? 1 : 0;   // more windows needed than the table holds
        }
    }
    __syncthreads();
    const int K = counts[2];
    {
        const bool too_many = counts[3] != 0;
        __syncthreads();   // as above: all reads of counts[] done before a return
        if (too_many) return false;
    }
    // window of every bucket: each thread takes B / nt consecutive buckets, finds the window
    // of the first one (last w with wb[w] <= b) and walks on (wb[K] = B ends the walk)
    {
        const int per = B / nt;
        const int b0 = tid * per;
        int lo = 0, hi = K;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (wb[mid] <= b0) lo = mid; else hi = mid;
        }
        int next = wb[lo + 1];
        for (int b = b0; b < b0 + per; ++b) {
            while (b >= next) next = wb[++lo + 1];
            wtab[b] = (uint8_t)lo;
        }
    }
    for (int w = tid; w < K; w += nt) wcur[w] = ws[w];
    if (tid == 0) counts[1] = 0;   // the first window's 17+-key bucket count
    __syncthreads();
    // one pass: every sample's index into its window's range of `binned`, its bucket folded
    // again (the times are shared by the curve's periods and stay in cache; a stored bucket
    // would be per-period scratch traffic)
    {
        constexpr int UW = 16;   // samples in flight per thread (20 or more spill into AGPRs)
        for (int64_t base = 0; base < n; base += (int64_t)UW * nt) {
            int bv[UW];
            double tq[UW];
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                const int64_t j = base + (int64_t)u * nt + tid;
                tq[u] = j < n ? t[j] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                const int64_t j = base + (int64_t)u * nt + tid;
                bv[u] = j < n ? phase_bucket(fold_phase_r(tq[u], P, rP), B, neg) : -1;
            }
            // consecutive lanes hold consecutive samples, whose windows rise within a cycle
            // of the period: each run of lanes with one window takes its slots with one
            // atomic of its first lane
            const int lane = lane_id();
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                const int64_t j = base + (int64_t)u * nt + tid;
                const int win = bv[u] >= 0 ? (int)wtab[bv[u]] : -1;
                const int prev = __shfl_up(win, 1, kWave);
                const bool head = lane == 0 || prev != win;
                const unsigned long long heads = __ballot(head);
                int first = 0;
                if (head && win >= 0) {
                    const unsigned long long above = lane == kWave - 1 ? 0ULL : heads >> (lane + 1);
                    const int len = above ? __ffsll((long long)above) : kWave - lane;
                    first = atomicAdd(&wcur[win], len);
                }
                const unsigned long long upto =
                    heads & (lane == kWave - 1 ? ~0ULL : ((2ULL << lane) - 1ULL));
                const int hl = 63 - __clzll((long long)upto);
                const int hfirst = __shfl(first, hl, kWave);
                if (win >= 0) binned[hfirst + (lane - hl)] = (int32_t)j;
            }
        }
    }
    __syncthreads();
    for (int w = 0; w < K; ++w) {
        const int bfirst = wb[w], bnext = wb[w + 1];
        const int wstart = ws[w], wcnt = ws[w + 1] - wstart;
        // stage the window's samples at their bucket positions (indices, then times, of
        // all of a thread's samples in flight before the first use)
        {
            int jj[kWinPer];
            double tv[kWinPer];
#pragma unroll
            for (int u = 0; u < kWinPer; ++u) {
                const int e = tid + u * nt;
                jj[u] = e < wcnt ? binned[wstart + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < kWinPer; ++u) tv[u] = t[jj[u] >= 0 ? jj[u] : 0];
#pragma unroll
            for (int u = 0; u < kWinPer; ++u) {
                if (jj[u] >= 0) {
                    const double ph = fold_phase_r(tv[u], P, rP);
                    const int pos = atomicAdd(&hist[hslot(phase_bucket(ph, B, neg))], 1) - wstart;
                    sphi[pos] = ph;
                    sidx[pos] = (int32_t)jj[u];
                }
            }
        }
        __syncthreads();
        // hist[b] now holds the end of bucket b for the window's buckets
        for (int b = bfirst + tid; b < bnext; b += nt) {
            const int blo = (b > bfirst ? hist[hslot(b - 1)] : wstart) - wstart;
            const int cnt = hist[hslot(b)] - wstart - blo;
            // one network per pass of the wavefront, the smallest that takes its largest
            // bucket of up to 16 keys (lanes would otherwise run several networks)
            const bool w16 = __ballot(cnt > 12 && cnt <= 16) != 0;
            const bool w12 = __ballot(cnt > 8 && cnt <= 12) != 0;
            if (cnt > 1 && cnt <= 16) {
                if (w16) sort_bucket_regs<16>(sphi, sidx, blo, cnt);
                else if (w12) sort_bucket_regs<12>(sphi, sidx, blo, cnt);
                else sort_bucket_regs<8>(sphi, sidx, blo, cnt);
            } else if (cnt > 16) {
                const int e = atomicAdd(&counts[1], 1);
                mids[e] = ((int64_t)blo << 32) | (int64_t)cnt;
            }
        }
        __syncthreads();
        {
            const int nmid = counts[1];
            const int wave = uniform((int32_t)wave_id());
            for (int e = wave; e < nmid; e += nt / kWave) {
                const int64_t ent = uniform(mids[e]);
                const int64_t blo = ent >> 32;
                const int cnt = (int)(ent & 0xffffffffLL);
                if (cnt <= 64) wave_bitonic<1>(sphi, sidx, blo, cnt);
                else if (cnt <= 128) wave_bitonic<2>(sphi, sidx, blo, cnt);
                else wave_bitonic<4>(sphi, sidx, blo, cnt);
            }
        }
        __syncthreads();
        // counts[1] is not read again in this window; the barrier after the sink orders the
        // reset before the next window's 17+-key buckets are counted
        if (tid == 0) counts[1] = 0;
        sink(wstart, wcnt);
        __syncthreads();
    }
    return true;
}

}  // namespace gtls
