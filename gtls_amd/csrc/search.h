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
#include "windows.h"

namespace gtls {

constexpr int kBlock = 256;
constexpr int64_t kTmplTail = 32768;  // zeroed samples after the last template row
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
    const Real* tmpl1;      // per template sample S
    const double* ss;       // per width: sum of S^2 over the row
    const double* chi2_flat;  // per curve
    const double* wconst;   // per curve: the common weight 1/dy^2, or NaN if dy varies
    int64_t skip_point;
    double depth_min;
    int32_t negative_time;
    int64_t total_periods;
    int32_t use_mfma;     // float32: window sums on the matrix cores
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
    __shared__ float rings[2][kBlock / kWave][kRingRows * kRingStride];

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

        // 2. gather in phase order (wrapping) and the float64 running sum of u. With a
        //    common weight only u is kept (X = u), otherwise the pairs (u w, w).
        const double wc = p.wconst[c];
        const bool constw = wc == wc;
        Real* ub = (Real*)vw;
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
                        if (constw) {
                            ub[i] = uu;
                        } else {
                            R2 o;
                            o.x = uu * ww;
                            o.y = ww;
                            vw[i] = o;
                        }
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
            // finite values past the sequence: the stride-1 window path multiplies them
            // by zero template padding
            if (tid < kTmplPad) {
                if (constw) ub[np_ + tid] = (Real)0;
                else vw[np_ + tid] = R2{(Real)0, (Real)0};
            }
        }
        __syncthreads();

        // 3. windows: wavefront w takes widths klo + w, klo + w + nwaves, ...; every
        //    per-width value is made wave-uniform so that template reads are scalar loads
        //    and the segment predicates are scalar branches
        const double chi2f = p.chi2_flat[c];
        const double sentinel = (double)n;
        ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};
        {
            const int lane = lane_id();
            const int wave = uniform((int32_t)wave_id());
            const int nwaves = kBlock / kWave;
            for (int32_t k = klo + wave; k <= khi; k += nwaves) {
                const int64_t wg = w0 + k;                      // global width index
                const int32_t d = uniform(p.widths[wg]);
                if (d <= 0) {  // empty window: mean 0/0, every position carries the sentinel
                    if (lane == 0) consider<Real>(best, sentinel, k, 0);
                    continue;
                }
                const int32_t q = (int32_t)((d > p.skip_point) ? (d / p.skip_point) : 1);
                const double os2 = 2.0 * p.overshoot[wg];
                const int64_t toff = uniform(p.tmpl_off[wg] + 2 * kTmplPad * wg + kTmplPad);
                if (q > 1 && n > 1 && lane == 0) consider<Real>(best, sentinel, k, 1);
                const int64_t K = (n + q - 1) / q;
                // candidates per lane: enough to cover K with one pass when possible, and
                // (R - 1) q within the slot's padding of max_width samples
                int32_t R = (int32_t)min((int64_t)kMaxBlock, max((int64_t)1, (K + kWave - 1) / kWave));
                R = uniform(min(R, 1 + p.max_width / q));
                const double ssk = p.ss[wg];
                const int32_t F = d / q;
                if constexpr (sizeof(Real) == 4) {
                    if (p.use_mfma && q >= 4 && F + 1 <= 16) {
                        // candidate k gets its window sums from the matrix-core tiles
                        auto emit = [&](int64_t kc, float sa, float sb, double dlo, double dhi) {
                            const int64_t s = kc * q;
                            const double m = (dhi - dlo) / (double)d;
                            if (m > p.depth_min) {
                                const float rs = (float)(m * os2);
                                const double delta =
                                    constw ? wc * ((double)(rs * rs) * ssk - 2.0 * (double)rs * (double)sa)
                                           : (double)(rs * rs * sb - 2.f * rs * sa);
                                consider<Real>(best, chi2f + delta, k, (int32_t)s);
                            } else {
                                consider<Real>(best, sentinel, k, (int32_t)s);
                            }
                        };
                        float* ra = rings[0][wave];
                        float* rb = rings[1][wave];
                        if (constw)
                            window_sums_mfma<true>((const float*)ub, np_, (const float*)(p.tmpl1 + toff),
                                                   D, d, q, K, ra, rb, emit);
                        else
                            window_sums_mfma<false>((const float*)vw, np_,
                                                    (const float*)(p.tmpl2 + toff), D, d, q, K, ra,
                                                    rb, emit);
                        continue;
                    }
                }
                for (int64_t base = 0; base < K; base += (int64_t)kWave * R) {
                    const int64_t k0 = base + (int64_t)lane * R;
                    const int32_t rc = (int32_t)max((int64_t)0, min((int64_t)R, K - k0));
                    Real A[kMaxBlock], B[kMaxBlock];
#pragma unroll
                    for (int r = 0; r < kMaxBlock; ++r) { A[r] = 0; B[r] = 0; }
                    const int64_t s0 = (rc > 0 ? k0 : 0) * q;
                    if (q == 1 && R == kMaxBlock) {
                        if (constw)
                            window_sums_q1<Real, true>(ub + s0, p.tmpl1 + toff, d, A, B);
                        else
                            window_sums_q1<Real, false>(vw + s0, (const Real*)(p.tmpl2 + toff), d, A, B);
                    } else {
                        if (constw)
                            window_sums_seg<Real, true>(ub + s0, p.tmpl1 + toff, d, q, R, A, B);
                        else
                            window_sums_seg<Real, false>(vw + s0, (const Real*)(p.tmpl2 + toff), d, q, R, A, B);
                    }
#pragma unroll
                    for (int r = 0; r < kMaxBlock; ++r) {
                        if (r < rc) {
                            const int64_t s = (k0 + r) * q;
                            const double m = (D[s + d - 1] - (s ? D[s - 1] : 0.0)) / (double)d;
                            if (m > p.depth_min) {
                                const double rr = m * os2;
                                double delta;
                                if (constw) {
                                    const Real rs = (Real)rr;
                                    delta = wc * ((double)(rs * rs) * ssk - 2.0 * (double)rs * (double)A[r]);
                                } else {
                                    const Real rs = (Real)rr;
                                    delta = (double)(rs * rs * B[r] - (Real)2 * rs * A[r]);
                                }
                                consider<Real>(best, chi2f + delta, k, (int32_t)s);
                            } else {
                                consider<Real>(best, sentinel, k, (int32_t)s);
                            }
                        }
                    }
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

// Search-precision template rows, (S, S^2) pairs and S alone, each row surrounded by
// kTmplPad zeros: row w starts at tmpl_off[w] + 2 kTmplPad w + kTmplPad. One block per row.
template <class Real>
__global__ void prep_template_kernel(const double* S, const int64_t* off, const int32_t* widths,
                                     int64_t nw, int64_t total_tmpl,
                                     typename Pair<Real>::type* out, Real* out1) {
    const int64_t w = blockIdx.x;
    if (w >= nw) {
        // the zeroed tail after the last row (over-reads of the matrix-core path)
        const int64_t end = total_tmpl + 2 * kTmplPad * nw;
        for (int64_t i = end + (int64_t)(w - nw) * blockDim.x + threadIdx.x;
             i < end + kTmplTail; i += (int64_t)(gridDim.x - nw) * blockDim.x) {
            typename Pair<Real>::type o;
            o.x = (Real)0;
            o.y = (Real)0;
            out[i] = o;
            out1[i] = (Real)0;
        }
        return;
    }
    const int32_t d = max(widths[w], 0);
    const int64_t src = off[w];
    const int64_t dst = off[w] + 2 * kTmplPad * w;   // start of the padded row
    for (int32_t i = threadIdx.x; i < d + 2 * kTmplPad; i += blockDim.x) {
        const int32_t j = i - kTmplPad;
        const double s = (j >= 0 && j < d) ? S[src + j] : 0.0;
        typename Pair<Real>::type o;
        o.x = (Real)s;
        o.y = (Real)(s * s);
        out[dst + i] = o;
        out1[dst + i] = (Real)s;
    }
}

// Sum of S^2 over each template row (the in-transit weight sum when all weights agree),
// with S rounded to the search precision first; one wavefront per width.
template <class Real>
__global__ __launch_bounds__(kWave) void template_ss_kernel(const double* S, const int64_t* off,
                                                           const int32_t* widths, int64_t nw,
                                                           double* ss) {
    const int64_t k = blockIdx.x;
    if (k >= nw) return;
    const int32_t d = widths[k];
    double acc = 0.0;
    for (int32_t i = threadIdx.x; i < d; i += kWave) {
        const double s = (double)(Real)S[off[k] + i];
        acc += s * s;
    }
    acc = wave_sum(acc);
    if (threadIdx.x == 0) ss[k] = acc;
}

// Flat-model chi2 of each curve, sum (1 - y)^2 / dy^2, float64, and the curve's common
// weight 1/dy^2 when every dy is the same value (NaN otherwise); one workgroup per curve.
__global__ __launch_bounds__(kBlock) void chi2_flat_kernel(const int64_t* pt_off,
                                                           const double* y, const double* dy,
                                                           double* out, double* wconst) {
    __shared__ double red[kBlock / kWave];
    __shared__ int differ;
    const int c = blockIdx.x;
    const int64_t a = pt_off[c], b = pt_off[c + 1];
    if (threadIdx.x == 0) differ = 0;
    __syncthreads();
    const double e0 = dy[a];
    double s = 0.0;
    int dif = 0;
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
        const double u = 1.0 - y[i];
        const double e = dy[i];
        const double w = 1.0 / (e * e);
        s += u * u * w;
        dif |= (e != e0);
    }
    if (dif) atomicOr(&differ, 1);
    s = wave_sum(s);
    if (lane_id() == 0) red[wave_id()] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) tot += red[w];
        out[c] = tot;
        wconst[c] = differ ? __builtin_nan("") : 1.0 / (e0 * e0);
    }
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ULL; }

}  // namespace gtls
