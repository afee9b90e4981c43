// The chi2-vs-period search of one (light curve, trial period) work item, executed by one
// workgroup of a persistent kernel, and the kernels around it.
//
// Per item (reference chain: src/gputls/core.py:770-875, GPUFun.cu:28-436):
//   1. fold + stable phase sort (fold_sort.h)                  -> perm
//   2. gather in phase order with the wrap-around extension    -> (v, w) pairs, prefix D
//      (patchData GPUFun.cu:108-127 with i -> perm[i mod N]); u = 1 - y, w = 1/dy^2,
//      v = u w; D = running sum of u in float64 (window means, GPUFun.cu:364-373)
//   3. every width k of the period, every candidate position s = 0, q, 2q, ... (q = skip
//      stride, GPUFun.cu:397): gate m > depth_min on the window mean m = (D[s+d-1] -
//      D[s-1]) / d, then the statistic
//         stat = chi2_flat + r^2 * sum_i S2_i w_{s+i} - 2 r * sum_i S_i v_{s+i},
//         r = 2 m overshoot_k,
//      which equals in-transit residual + out-of-transit residual - edge correction of
//      GPUFun.cu:380-436 (the out-of-transit part and the edge correction cancel to the
//      order-free flat-model chi2 of the unpatched samples, chi2_flat = sum (1-y)^2/dy^2).
//      Positions that are not evaluated carry the statistic N (GPUFun.cu:400).
//   4. first minimum in (width, position) order (cp.argmin, core.py:866-875).
#pragma once
#include "fold_sort.h"

namespace gtls {

constexpr int kBlock = 256;
constexpr int kScanPerThread = 8;  // gathered elements per thread per prefix-scan chunk

template <class Real> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

template <class Real>
struct SearchParams {
    using R2 = typename Pair<Real>::type;
    int32_t n_curves;
    const int64_t* pt_off;
    const double* t;
    const int64_t* per_off;
    const double* periods;
    const int32_t* k_lo;
    const int32_t* k_hi;
    const int64_t* wid_off;
    const int32_t* widths;
    const double* overshoot;
    const int64_t* tmpl_off;
    const R2* uw0;          // per point (u, w)
    const R2* tmpl2;        // per template sample (S, S^2)
    const double* chi2_flat;  // per curve
    int64_t skip_point;
    double depth_min;
    int32_t negative_time;
    int64_t total_periods;
    int32_t max_points;   // slot capacity: every curve must have n <= max_points
    int32_t max_width;    // and every searched width d <= max_width
    unsigned long long* queue;
    char* slots;
    int64_t slot_bytes;
    int64_t o_key_in, o_key_out, o_perm, o_vw, o_D;
    double* chi2;
    int32_t* best_k;
    int32_t* best_s;
};

// Largest c with off[c] <= g (off ascending, off[0] = 0); wave-uniform.
__device__ __forceinline__ int find_owner(const int64_t* off, int n, int64_t g) {
    int lo = 0, hi = n;  // invariant: off[lo] <= g < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

template <class Real>
__device__ __forceinline__ void consider(ArgMin& b, double v, int32_t k, int32_t s) {
    ArgMin c{v, k, s};
    if (argmin_less(c, b)) b = c;
}

template <class Real>
__global__ __launch_bounds__(kBlock) void tls_search_kernel(SearchParams<Real> p) {
    using R2 = typename Pair<Real>::type;
    __shared__ int hist[kMaxBuckets];
    __shared__ int red_i[kBlock / kWave];
    __shared__ double red_d[kBlock / kWave];
    __shared__ ArgMin red_a[kBlock / kWave];
    __shared__ unsigned long long s_item;

    const int tid = threadIdx.x;
    char* slot = p.slots + (int64_t)blockIdx.x * p.slot_bytes;
    double* key_in = (double*)(slot + p.o_key_in);
    double* key_out = (double*)(slot + p.o_key_out);
    int32_t* perm = (int32_t*)(slot + p.o_perm);
    R2* vw = (R2*)(slot + p.o_vw);
    double* D = (double*)(slot + p.o_D);

    for (;;) {
        if (tid == 0) s_item = atomicAdd(p.queue, 1ULL);
        __syncthreads();
        const int64_t g = uniform((int64_t)s_item);
        __syncthreads();
        if (g >= p.total_periods) break;

        // Every value below derives from g and is wave-uniform; the item either runs all
        // three stages or only writes its status, and both paths meet at the barrier
        // that ends the iteration (no early continue out of a loop with barriers).
        const int c = find_owner(p.per_off, p.n_curves, g);
        const int64_t pt0 = p.pt_off[c];
        const int64_t n = p.pt_off[c + 1] - pt0;
        const int64_t w0 = p.wid_off[c];
        const int32_t klo = p.k_lo[g], khi = p.k_hi[g];
        // status: 0 run, -1 no widths (NaN), -2 inputs that would overrun the slot
        int32_t status = (klo > khi || n <= 0) ? -1 : 0;
        int32_t dmax = 1;
        if (status == 0) {
            const int64_t nw = p.wid_off[c + 1] - w0;
            if (klo < 0 || khi >= nw || n > p.max_points) {
                status = -2;
            } else {
                for (int32_t k = klo; k <= khi; ++k) dmax = max(dmax, p.widths[w0 + k]);
                if (dmax > p.max_width) status = -2;
            }
        }
        if (status != 0) {
            if (tid == 0) {
                p.chi2[g] = __builtin_nan("");
                p.best_k[g] = status;
                p.best_s[g] = status;
            }
        } else {
        const double P = p.periods[g];

        // 1. fold + sort
        wg_fold_sort(p.t + pt0, n, P, p.negative_time != 0, key_in, key_out, perm, hist,
                     red_i);

        // 2. gather in phase order (wrapping) and the float64 running sum of u
        const int64_t np_ = n + dmax;
        {
            const int lane = lane_id(), wave = wave_id();
            const int64_t chunk = (int64_t)kBlock * kScanPerThread;
            double carry = 0.0;
            for (int64_t base = 0; base < np_; base += chunk) {
                const int64_t wbase = base + (int64_t)wave * kWave * kScanPerThread;
                double u[kScanPerThread];
                double run = 0.0;
#pragma unroll
                for (int e = 0; e < kScanPerThread; ++e) {
                    const int64_t i = wbase + (int64_t)e * kWave + lane;
                    Real uu = 0, ww = 0;
                    if (i < np_) {
                        const int64_t j = i < n ? i : i % n;
                        const R2 a = p.uw0[pt0 + perm[j]];
                        uu = a.x;
                        ww = a.y;
                        R2 o;
                        o.x = uu * ww;
                        o.y = ww;
                        vw[i] = o;
                    }
                    const double s = wave_inclusive_scan((double)uu);
                    u[e] = s + run;
                    run += __shfl(s, kWave - 1, kWave);
                }
                if (lane == kWave - 1) red_d[wave] = run;
                __syncthreads();
                double before = carry, total = carry;
                for (int w = 0; w < kBlock / kWave; ++w) {
                    if (w < wave) before += red_d[w];
                    total += red_d[w];
                }
                __syncthreads();
#pragma unroll
                for (int e = 0; e < kScanPerThread; ++e) {
                    const int64_t i = wbase + (int64_t)e * kWave + lane;
                    if (i < np_) D[i] = u[e] + before;
                }
                carry = total;
            }
        }
        __syncthreads();

        // 3. windows of every width of the period
        const double chi2f = p.chi2_flat[c];
        const double sentinel = (double)n;
        ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};
        for (int32_t k = klo; k <= khi; ++k) {
            const int32_t d = p.widths[w0 + k];
            if (d <= 0) {  // empty window: mean 0/0, every position carries the sentinel
                if (tid == 0) consider<Real>(best, sentinel, k, 0);
                continue;
            }
            const int64_t q = (d > p.skip_point) ? (d / p.skip_point) : 1;
            const double os2 = 2.0 * p.overshoot[w0 + k];
            const R2* T = p.tmpl2 + p.tmpl_off[w0 + k];
            if (q > 1 && n > 1 && tid == 0) consider<Real>(best, sentinel, k, 1);
            const int64_t ncand = (n + q - 1) / q;
            for (int64_t ci = tid; ci < ncand; ci += kBlock) {
                const int64_t s = ci * q;
                const double m = (D[s + d - 1] - (s ? D[s - 1] : 0.0)) / (double)d;
                if (m > p.depth_min) {
                    const Real r = (Real)(m * os2);
                    Real A = 0, Bw = 0;
                    const R2* x = vw + s;
                    for (int32_t i = 0; i < d; ++i) {
                        const R2 xv = x[i];
                        const R2 tv = T[i];
                        A = fma(tv.x, xv.x, A);
                        Bw = fma(tv.y, xv.y, Bw);
                    }
                    const Real delta = r * r * Bw - (Real)2 * r * A;
                    consider<Real>(best, chi2f + (double)delta, k, (int32_t)s);
                } else {
                    consider<Real>(best, sentinel, k, (int32_t)s);
                }
            }
        }
        best = block_argmin(best, red_a);
        if (tid == 0) {
            const bool found = best.k != 0x7fffffff;
            p.chi2[g] = found ? best.v : __builtin_nan("");
            p.best_k[g] = found ? best.k : -1;
            p.best_s[g] = found ? best.s : -1;
        }
        }  // status == 0
        __syncthreads();
    }
}

// Stable phase permutations only (bit-exact checks of the sort).
__global__ __launch_bounds__(kBlock) void fold_sort_kernel(const double* t, int64_t n,
                                                          const double* periods,
                                                          int64_t n_periods, int neg,
                                                          char* slots, int64_t slot_bytes,
                                                          int32_t* perm_out) {
    __shared__ int hist[kMaxBuckets];
    __shared__ int red_i[kBlock / kWave];
    char* slot = slots + (int64_t)blockIdx.x * slot_bytes;
    double* key_in = (double*)slot;
    double* key_out = key_in + n;
    for (int64_t pi = blockIdx.x; pi < n_periods; pi += gridDim.x)
        wg_fold_sort(t, n, periods[pi], neg != 0, key_in, key_out, perm_out + pi * n, hist,
                     red_i);
}

template <class Real>
__global__ void prep_points_kernel(const double* y, const double* dy, int64_t n,
                                   typename Pair<Real>::type* uw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = 1.0 - y[i];
    const double e = dy[i];
    const double w = 1.0 / (e * e);
    typename Pair<Real>::type o;
    o.x = (Real)u;
    o.y = (Real)w;
    uw[i] = o;
}

template <class Real>
__global__ void prep_template_kernel(const double* S, int64_t n, typename Pair<Real>::type* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = S[i];
    typename Pair<Real>::type o;
    o.x = (Real)s;
    o.y = (Real)(s * s);
    out[i] = o;
}

// Flat-model chi2 of each curve, sum (1 - y)^2 / dy^2, float64; one workgroup per curve.
__global__ __launch_bounds__(kBlock) void chi2_flat_kernel(const int64_t* pt_off,
                                                           const double* y, const double* dy,
                                                           double* out) {
    __shared__ double red[kBlock / kWave];
    const int c = blockIdx.x;
    const int64_t a = pt_off[c], b = pt_off[c + 1];
    double s = 0.0;
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
        const double u = 1.0 - y[i];
        const double e = dy[i];
        const double w = 1.0 / (e * e);
        s += u * u * w;
    }
    s = wave_sum(s);
    if (lane_id() == 0) red[wave_id()] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) tot += red[w];
        out[c] = tot;
    }
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ULL; }

}  // namespace gtls
