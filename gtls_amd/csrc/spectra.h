// Signal residue, SDE and detrended power of every light curve of a batch, from its
// chi2-vs-period vector (src/gputls/stats.py:95-122, running median of
// src/gputls/helpers.py:91-106), in float64.
//
//   pass 1 (workgroup per curve): min, SR = min/chi2, SDE_raw = (1 - mean SR) / std SR,
//                                 power_raw = (SR - mean) * SDE_raw / max(SR - mean)
//   median (thread per period):   91-wide (3 * 30 + 1) running median of power_raw, the
//                                 exact middle order statistic of each window, edges
//                                 padded with the first / last full window
//   pass 2 (workgroup per curve): power = power_raw - median, minus its mean;
//                                 SDE = max(power / std); power *= SDE / max(power)
//
// Curves whose chi2 holds a NaN, whose largest chi2 exceeds 100x its smallest (the only
// case in which the reference's outlier mask, src/gputls/core.py:890-896, can be
// non-empty), or whose SR or detrended power has zero spread (NumPy's masked division
// then masks the result) get status 1 and are left to the host.
#pragma once
#include "wg_primitives.h"

namespace gtls {

constexpr int kSpecBlock = 256;

struct SpecParams {
    int32_t n_curves;
    const int64_t* per_off;
    const double* chi2;
    int32_t kernel;      // odd running-median width
    double* power_raw;   // [total]
    double* power;       // [total]
    double* sde_raw;     // [n_curves]
    double* sde;         // [n_curves]
    int32_t* status;     // [n_curves]: 0 done on device, 1 host fallback
};

__device__ __forceinline__ double block_reduce_sum(double v, double* tmp) {
    v = wave_sum(v);
    if (lane_id() == 0) tmp[wave_id()] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += tmp[w];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double block_reduce_max(double v, double* tmp) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    if (lane_id() == 0) tmp[wave_id()] = v;
    __syncthreads();
    double s = tmp[0];
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) s = fmax(s, tmp[w]);
    __syncthreads();
    return s;
}

__device__ __forceinline__ double block_reduce_min(double v, double* tmp) {
    return -block_reduce_max(-v, tmp);
}

__global__ __launch_bounds__(kSpecBlock) void spectra_pass1_kernel(SpecParams p) {
    __shared__ double tmp[kSpecBlock / kWave];
    const int c = blockIdx.x;
    const int64_t a = p.per_off[c], b = p.per_off[c + 1];
    const int64_t n = b - a;
    const int tid = threadIdx.x;
    if (n <= 0) {
        if (tid == 0) { p.status[c] = 1; }
        return;
    }
    double lo = __builtin_inf(), hi = -__builtin_inf(), nans = 0.0;
    for (int64_t i = a + tid; i < b; i += blockDim.x) {
        const double x = p.chi2[i];
        if (x != x) nans += 1.0;
        else { lo = fmin(lo, x); hi = fmax(hi, x); }
    }
    lo = uniform(block_reduce_min(lo, tmp));
    hi = uniform(block_reduce_max(hi, tmp));
    nans = uniform(block_reduce_sum(nans, tmp));
    if (nans > 0.0 || !(hi <= 100.0 * lo)) {
        if (tid == 0) p.status[c] = 1;
        return;
    }
    // SR mean and (two-pass) standard deviation
    double s = 0.0;
    for (int64_t i = a + tid; i < b; i += blockDim.x) s += lo / p.chi2[i];
    const double mean = block_reduce_sum(s, tmp) / (double)n;
    double v = 0.0, mx = -__builtin_inf();
    for (int64_t i = a + tid; i < b; i += blockDim.x) {
        const double d = lo / p.chi2[i] - mean;
        v += d * d;
        mx = fmax(mx, d);
    }
    const double sd = uniform(sqrt(block_reduce_sum(v, tmp) / (double)n));
    mx = uniform(block_reduce_max(mx, tmp));
    if (!(sd > 0.0) || !(mx > 0.0)) {  // zero spread: the masked-array result is the host's
        if (tid == 0) p.status[c] = 1;
        return;
    }
    const double sde_raw = (1.0 - mean) / sd;
    const double scale = sde_raw / mx;
    const bool detrend = n > 2 * (int64_t)p.kernel;
    for (int64_t i = a + tid; i < b; i += blockDim.x) {
        const double pr = (lo / p.chi2[i] - mean) * scale;
        p.power_raw[i] = pr;
        if (!detrend) p.power[i] = pr;
    }
    if (tid == 0) {
        p.sde_raw[c] = sde_raw;
        p.status[c] = 0;
        if (!detrend) p.sde[c] = sde_raw;
    }
}

// Running median, one thread per period of the whole batch. The window of output i of a
// curve starts at clamp(i - front, 0, n - k) inside that curve, so it lies within
// [i - k, i + k] of the global index; a block stages that halo of its 256 outputs in LDS
// and each thread ranks the values of its window (exact middle order statistic).
constexpr int kMaxMedianKernel = 255;

__device__ __forceinline__ int owner_curve(const int64_t* off, int n, int64_t g) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kSpecBlock) void running_median_kernel(SpecParams p,
                                                                    int64_t total,
                                                                    double* med) {
    __shared__ double win[kSpecBlock + 2 * kMaxMedianKernel + 2];
    const int k = p.kernel;
    const int64_t I0 = (int64_t)blockIdx.x * kSpecBlock;
    const int64_t L0 = max(I0 - (int64_t)k, (int64_t)0);
    const int64_t L1 = min(I0 + kSpecBlock + k, total);
    for (int64_t e = L0 + threadIdx.x; e < L1; e += blockDim.x) win[e - L0] = p.power_raw[e];
    __syncthreads();
    const int64_t i = I0 + threadIdx.x;
    if (i >= total) return;
    const int c = owner_curve(p.per_off, p.n_curves, i);
    const int64_t a = p.per_off[c], n = p.per_off[c + 1] - a;
    if (p.status[c] != 0 || n <= 2 * (int64_t)k) return;
    const int64_t nwin = n - k + 1;
    const int64_t front = (n - nwin) / 2;
    const int64_t j = min(max((i - a) - front, (int64_t)0), nwin - 1);
    const double* w = win + (a + j - L0);
    const int m = k / 2;
    double res = w[0];
    for (int ci = 0; ci < k; ++ci) {
        const double vc = w[ci];
        int less = 0, leq = 0;
        for (int e = 0; e < k; ++e) {
            const double x = w[e];
            less += x < vc;
            leq += x <= vc;
        }
        if (less <= m && m < leq) { res = vc; break; }
    }
    med[i] = res;
}

__global__ __launch_bounds__(kSpecBlock) void spectra_pass2_kernel(SpecParams p,
                                                                   const double* med) {
    __shared__ double tmp[kSpecBlock / kWave];
    const int c = blockIdx.x;
    const int64_t a = p.per_off[c], b = p.per_off[c + 1];
    const int64_t n = b - a;
    if (p.status[c] != 0 || n <= 2 * (int64_t)p.kernel) return;
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int64_t i = a + tid; i < b; i += blockDim.x) s += p.power_raw[i] - med[i];
    const double m1 = block_reduce_sum(s, tmp) / (double)n;
    double s2 = 0.0;
    for (int64_t i = a + tid; i < b; i += blockDim.x) s2 += (p.power_raw[i] - med[i]) - m1;
    const double m2 = block_reduce_sum(s2, tmp) / (double)n;
    double v = 0.0, mx = -__builtin_inf();
    for (int64_t i = a + tid; i < b; i += blockDim.x) {
        const double q = (p.power_raw[i] - med[i]) - m1;
        const double d = q - m2;
        v += d * d;
        mx = fmax(mx, q);
    }
    const double sd = uniform(sqrt(block_reduce_sum(v, tmp) / (double)n));
    mx = uniform(block_reduce_max(mx, tmp));
    if (!(sd > 0.0) || !(mx > 0.0)) {
        if (tid == 0) p.status[c] = 1;
        return;
    }
    const double sde = mx / sd;
    const double scale = sde / mx;
    for (int64_t i = a + tid; i < b; i += blockDim.x)
        p.power[i] = ((p.power_raw[i] - med[i]) - m1) * scale;
    if (tid == 0) p.sde[c] = sde;
}

}  // namespace gtls
