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
constexpr int64_t kTmplTail = 32768;  // least zeroed samples after the last template row
constexpr int kScanPerThread = 11;  // gathered elements per thread per prefix-scan chunk
// Queue counters sit 256 bytes apart (no two in one cache line: the atomics of different
// XCDs on one line serialise); kChunkCounters per period chunk.
constexpr int kCounterPitch = 32;
constexpr int kChunkCounters = 16 * kCounterPitch;
// Window items taken per queue atomic. With 1, consecutive items (the widths of one period)
// go to different wavefronts, which then stream the period's samples through L1 and L2
// together (measured: 2 items +0.6 %, 4 items +6 %, 8 items +18 % kernel time).
constexpr int kItemBatch = 1;
static_assert(kBlock * kScanPerThread >= kWinCap, "one prefix-scan chunk per sort window");


// Per trial period of a chunk: written by prep_kernel, read by the two later stages.
struct PeriodMeta {
    int32_t c;        // curve
    int32_t status;   // 0 searched, -1 no widths / no points (NaN), -2 inputs overrun a slot
    int64_t np;       // gathered samples: N + widest searched width
};

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
    const Real* tmplsq;     // per template sample S^2
    // float32 matrix-core operands in float16 (hi, lo) pair planes (windows.h):
    //   tmplh: template rows with kTmplPad zeros on both sides, a plane of S pairs and,
    //          tlo bytes after it, one of S^2 pairs (the Toeplitz tiles);
    //   tmplc: chunked template rows of the diagonal-sum tiles in the same two planes,
    //          clo bytes apart, each ending in a zero block at chunk offset zoff
    const char* tmplh;
    const char* tmplc;
    int32_t tlo, clo;
    int64_t zoff;
    const float2* xsc;      // per curve: power-of-two scales of the A-sum and weight planes
    const uint32_t* upair;  // float32: per point the (hi, lo) pair of u scaled by its curve's
                            // A-sum plane scale (the samples of a common weight)
    int32_t d_constw;       // float32: curves with a common weight need the prefix sums D
                            // (some width takes a path that reads them)
    const double* ss;       // per width: sum of S^2 over the row
    const double* chi2_flat;  // per curve
    const double* wconst;   // per curve: the common weight 1/dy^2, or NaN if dy varies
    int64_t skip_point;
    double depth_min;
    int32_t negative_time;
    int32_t use_mfma;     // float32: window sums on the matrix cores
    int32_t toep_q;       // float32: strides below this take the Toeplitz tiles (<= kToepQ)
    int32_t quad_q;       // float32, common weight: diagonal-sum strides from this on load
                          // their sample operands in the Quad layout (windows.h)
    int32_t max_points;   // every curve must have n <= max_points
    int32_t max_width;    // every searched width must be <= max_width
    int32_t kmax;         // items per period in the window stage (>= widths per period)
    // the chunk: periods [p0, p0 + chunk) of the batch
    int64_t p0;
    int64_t chunk;
    // prep_kernel's per-workgroup sort scratch
    char* slots;
    int64_t slot_bytes;
    int64_t o_key, o_perm, o_tmp_key, o_tmp_id, o_lists;
    // per chunk period i: gathered samples in xbuf[i * xstride ...] (room for 2 xstride
    // Reals: u, or the planes v = u w at [0, xstride) and w at [xstride, 2 xstride)) and
    // their float64 running sum dbuf[i * xstride ...]
    R2* xbuf;
    double* dbuf;
    int64_t xstride;
    PeriodMeta* meta;
    ArgMin* part;                  // per (chunk period, width slot): the width's minimum
    unsigned long long* counters;  // queues, kCounterPitch apart: the prep queue, then the
                                   // window queues (one per XCD)
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


// The XCD (accelerator complex die) this wavefront runs on, 0..7.
__device__ __forceinline__ int xcd_id() {
    int v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(v));
    return v & 7;
}

// Gathers the samples of sorted positions o0 .. o0 + cnt - 1 (source sample src(j) for
// j < cnt) and hands each to put(position, u, w), and writes the float64 running sum D of
// u (unless D is NULL), continuing from carry (updated; equal in every thread). All
// threads call it; red_d is LDS of blockDim/64.
template <class Real, class Src, class Put>
__device__ __forceinline__ void gather_scan(const typename Pair<Real>::type* __restrict__ uw,
                                            int64_t o0, int64_t cnt, Src src, Put put,
                                            double* D, double& carry, double* red_d) {
    using R2 = typename Pair<Real>::type;
    const int lane = lane_id(), wave = wave_id();
    const int64_t chunk = (int64_t)kBlock * kScanPerThread;
    for (int64_t base = 0; base < cnt; base += chunk) {
        const int64_t wbase = base + (int64_t)wave * kWave * kScanPerThread;
        double u[kScanPerThread];
        double run = 0.0;
        // all gathers of the chunk in flight before the first use
        R2 g[kScanPerThread];
#pragma unroll
        for (int e = 0; e < kScanPerThread; ++e) {
            const int64_t j = wbase + (int64_t)e * kWave + lane;
            g[e] = uw[src(j < cnt ? j : 0)];
        }
#pragma unroll
        for (int e = 0; e < kScanPerThread; ++e) {
            const int64_t j = wbase + (int64_t)e * kWave + lane;
            Real uu = 0;
            if (j < cnt) {
                uu = g[e].x;
                put(o0 + j, uu, g[e].y);
            }
            const double sc = wave_inclusive_scan((double)uu);
            u[e] = sc + run;
            run += __shfl(sc, kWave - 1, kWave);
        }
        if (lane == kWave - 1) red_d[wave] = run;
        __syncthreads();
        double before = carry, total = carry;
        for (int w = 0; w < kBlock / kWave; ++w) {
            if (w < wave) before += red_d[w];
            total += red_d[w];
        }
        __syncthreads();
        if (D) {
#pragma unroll
            for (int e = 0; e < kScanPerThread; ++e) {
                const int64_t j = wbase + (int64_t)e * kWave + lane;
                if (j < cnt) D[o0 + j] = u[e] + before;
            }
        }
        carry = total;
    }
}

// Where prep_kernel puts the gathered samples of one period (windows.h for the layouts):
// float64 -- u alone with a common weight, else the planes v = u w and w, at ub and wpl;
// float32 -- the float16 (hi, lo) pair planes of sx u (common weight) or of sx v and sw w,
// the first at hb and the second ps pairs after it.
template <class Real> struct SampleSink;
template <>
struct SampleSink<double> {
    double* ub;
    double* wpl;
    bool constw;
    __device__ __forceinline__ void put(int64_t j, double u, double w) const {
        if (constw) {
            ub[j] = u;
        } else {
            ub[j] = u * w;
            wpl[j] = w;
        }
    }
    // position j repeats position jj (the wrap-around extension) / is zero
    __device__ __forceinline__ void copy(int64_t j, int64_t jj) const {
        ub[j] = ub[jj];
        if (!constw) wpl[j] = wpl[jj];
    }
    __device__ __forceinline__ void zero(int64_t j) const {
        ub[j] = 0.0;
        if (!constw) wpl[j] = 0.0;
    }
};
template <>
struct SampleSink<float> {
    uint32_t* hb;
    int64_t ps;
    float sx, sw;
    bool constw;
    __device__ __forceinline__ void put(int64_t j, float u, float w) const {
        if (constw) {
            hb[j] = split_pair(u * sx);
        } else {
            hb[j] = split_pair((u * w) * sx);
            hb[j + ps] = split_pair(w * sw);
        }
    }
    __device__ __forceinline__ void copy(int64_t j, int64_t jj) const {
        hb[j] = hb[jj];
        if (!constw) hb[j + ps] = hb[jj + ps];
    }
    __device__ __forceinline__ void zero(int64_t j) const {
        hb[j] = 0u;
        if (!constw) hb[j + ps] = 0u;
    }
};

// gather_scan for one sort window of cnt <= kWinCap samples in float32 with a common
// weight: each thread gathers kScanPerThread consecutive positions, sums them serially,
// and one workgroup scan of the thread totals gives every position's running sum of u
// (float64); u and D are staged in the window's LDS (su over the sample indices sidx, sD
// over the phases) and copied out coalesced, u as the float16 (hi, lo) pairs of sx u at hb
// and D unless it is NULL. All threads call it.
__device__ __forceinline__ void gather_scan_window(const float2* __restrict__ uw, int64_t o0,
                                                   int cnt, int32_t* sidx, double* sD,
                                                   uint32_t* hb, float sx, double* D,
                                                   double& carry, double* red_d) {
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    const int p0 = tid * kScanPerThread;
    float* su = (float*)sidx;   // position p is read (as an index) before it is rewritten
    float u[kScanPerThread];
#pragma unroll
    for (int e = 0; e < kScanPerThread; ++e) {
        const int p = p0 + e;
        u[e] = p < cnt ? uw[sidx[p]].x : 0.f;
    }
    double pre[kScanPerThread];
    double run = 0.0;
#pragma unroll
    for (int e = 0; e < kScanPerThread; ++e) {
        run += (double)u[e];
        pre[e] = run;
    }
    const double incl = wave_inclusive_scan(run);
    if (lane == kWave - 1) red_d[wave] = incl;
    __syncthreads();
    double before = carry + (incl - run), total = carry;
    for (int w = 0; w < kBlock / kWave; ++w) {
        if (w < wave) before += red_d[w];
        total += red_d[w];
    }
#pragma unroll
    for (int e = 0; e < kScanPerThread; ++e) {
        const int p = p0 + e;
        if (p < cnt) {
            sD[p] = before + pre[e];
            su[p] = u[e];
        }
    }
    __syncthreads();
    // the windows kernel reads these only after the whole chunk is prepared: stored past
    // the caches, so that they do not evict the curves' times and sort scratch
    if (D) {
        for (int j = tid; j < cnt; j += kBlock) {
            __builtin_nontemporal_store(split_pair(su[j] * sx), &hb[o0 + j]);
            __builtin_nontemporal_store(sD[j], &D[o0 + j]);
        }
    } else {
        for (int j = tid; j < cnt; j += kBlock)
            __builtin_nontemporal_store(split_pair(su[j] * sx), &hb[o0 + j]);
    }
    carry = total;
}

// The samples of one sort window in float32 with a common weight when no running sums are
// wanted: position o0 + j takes the (hi, lo) pair of sx u of sample sidx[j], prepared per
// point in up (pair_points_kernel). All of a thread's gathers are in flight before its
// first store. All threads call it; no barrier.
__device__ __forceinline__ void gather_window_pairs(const uint32_t* __restrict__ up, int64_t o0,
                                                    int cnt, const int32_t* sidx, uint32_t* hb,
                                                    int64_t n, int64_t wrap) {
    uint32_t u[kWinPer];
#pragma unroll
    for (int e = 0; e < kWinPer; ++e) {
        const int j = threadIdx.x + e * kBlock;
        u[e] = j < cnt ? up[sidx[j]] : 0u;
    }
#pragma unroll
    for (int e = 0; e < kWinPer; ++e) {
        const int j = threadIdx.x + e * kBlock;
        if (j < cnt) {
            const uint32_t v = u[e];
            __builtin_nontemporal_store(v, &hb[o0 + j]);
            // the wrap-around extension: position n + p repeats position p < wrap
            if (o0 + j < wrap) __builtin_nontemporal_store(v, &hb[n + o0 + j]);
        }
    }
}

// Stage 1 + 2 of each period of the chunk, one workgroup per period (persistent):
//   1. fold + stable phase sort (fold_sort.h)                  -> perm
//   2. gather in phase order with the wrap-around extension    -> X, running sum D
// Two workgroups per CU (their LDS allows no more) at 256 registers: two waves per SIMD.
template <class Real>
__global__ __launch_bounds__(kBlock) void prep_kernel(SearchParams<Real> p) {
    using R2 = typename Pair<Real>::type;
    __shared__ __attribute__((aligned(16))) int hist[kHistLen];
    __shared__ __attribute__((aligned(16))) double s_phi[kWinCap];
    __shared__ int32_t s_idx[kWinCap];
    __shared__ int64_t s_mids[kWinMid];
    __shared__ uint8_t s_wtab[kMaxBuckets];
    __shared__ int s_wb[kMaxWin + 1], s_ws[kMaxWin + 1], s_wcur[kMaxWin];
    __shared__ int red_i[kBlock / kWave];
    __shared__ int s_counts[4];
    __shared__ double red_d[kBlock / kWave];
    __shared__ unsigned long long s_item;

    const int tid = threadIdx.x;
    char* slot = p.slots + (int64_t)blockIdx.x * p.slot_bytes;
    double* key = (double*)(slot + p.o_key);
    int32_t* perm = (int32_t*)(slot + p.o_perm);
    double* tmp_key = (double*)(slot + p.o_tmp_key);
    int32_t* tmp_id = (int32_t*)(slot + p.o_tmp_id);
    int64_t* lists = (int64_t*)(slot + p.o_lists);

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&p.counters[0], 1ULL);
        __syncthreads();
        const int64_t i = uniform((int64_t)s_item);
        __syncthreads();
        if (i >= p.chunk) break;
        const int64_t g = p.p0 + i;

        // Every value below derives from g and is wave-uniform; the item either runs both
        // stages or only writes its status, and both paths meet at the barrier that ends
        // the iteration (no early continue out of a loop with barriers).
        const int c = find_owner(p.per_off, p.n_curves, g);
        const int64_t pt0 = p.pt_off[c];
        const int64_t n = p.pt_off[c + 1] - pt0;
        const int64_t w0 = p.wid_off[c];
        const int32_t klo = p.k_lo[g], khi = p.k_hi[g];
        int32_t status = (klo > khi || n <= 0) ? -1 : 0;
        int32_t dmax = 1;
        if (status == 0) {
            const int64_t nw = p.wid_off[c + 1] - w0;
            if (klo < 0 || khi >= nw || n > p.max_points || khi - klo + 1 > p.kmax) {
                status = -2;
            } else {
                for (int32_t k = klo; k <= khi; ++k) dmax = max(dmax, p.widths[w0 + k]);
                if (dmax > p.max_width) status = -2;
            }
        }
        const int64_t np_ = n + dmax;
        if (tid == 0) p.meta[i] = PeriodMeta{c, status, np_};
        if (status == 0) {
            const double P = p.periods[g];
            // gather in phase order (wrapping) and the float64 running sum of u. With a
            // common weight only u is kept (X = u), otherwise the planes v = u w and w.
            const double wc = p.wconst[c];
            const bool constw = wc == wc;
            SampleSink<Real> sink;
            if constexpr (sizeof(Real) == 8) {
                sink.ub = (double*)(p.xbuf + i * p.xstride);
                sink.wpl = sink.ub + p.xstride;
            } else {
                const float2 sc = p.xsc[c];
                sink.hb = (uint32_t*)(p.xbuf + i * p.xstride);
                sink.ps = p.xstride;
                sink.sx = sc.x;
                sink.sw = sc.y;
            }
            sink.constw = constw;
            // float32 with a common weight: the running sums only if some width reads them
            double* D = (sizeof(Real) == 4 && constw && !p.d_constw) ? nullptr
                                                                    : p.dbuf + i * p.xstride;
            const R2* uw = p.uw0 + pt0;
            double carry = 0.0;
            auto put = [&](int64_t j, Real u, Real w) { sink.put(j, u, w); };
            // float32 with a common weight and no running sums: the windowed sink writes the
            // wrap-around copy itself when it needs one lap (dmax <= n)
            const int64_t wrap_in_sink =
                (sizeof(Real) == 4 && constw && !D && dmax <= n) ? (int64_t)dmax : 0;
            const bool windowed = wg_fold_sort_windows(
                p.t + pt0, n, P, p.negative_time != 0, perm, hist, red_i, s_counts, s_phi,
                s_idx, s_mids, s_wtab, s_wb, s_ws, s_wcur, [&](int wstart, int wcnt) {
                    if constexpr (sizeof(Real) == 4) {
                        if (constw && !D) {
                            gather_window_pairs(p.upair + pt0, wstart, wcnt, s_idx, sink.hb, n,
                                                wrap_in_sink);
                            return;
                        }
                        if (constw) {
                            gather_scan_window((const float2*)uw, wstart, wcnt, s_idx, s_phi,
                                               sink.hb, sink.sx, D, carry, red_d);
                            return;
                        }
                    }
                    gather_scan<Real>(uw, wstart, wcnt, [&](int64_t j) { return s_idx[j]; },
                                      put, D, carry, red_d);
                });
            if (!windowed) {
                // a bucket too large for LDS windows: the global form, then one gather pass
                wg_fold_sort(p.t + pt0, n, P, p.negative_time != 0, key, perm, tmp_key, tmp_id,
                             lists, hist, red_i, s_counts);
                gather_scan<Real>(uw, 0, n, [&](int64_t j) { return perm[j]; }, put, D, carry,
                                  red_d);
            }
            __syncthreads();
            // the wrap-around extension: positions n .. np - 1 repeat positions 0 .. dmax - 1
            // (already written by the sink when it took it; not when the global form ran)
            for (int64_t j = n + tid + ((windowed && wrap_in_sink) ? np_ - n : 0); j < np_; j += kBlock) {
                const int64_t jj = (j - n) % n;
                const int64_t laps = (j - n) / n + 1;
                sink.copy(j, jj);
                if (D) D[j] = (double)laps * carry + D[jj];
            }
            // zeros past the sequence, up to where the window loops read (below
            // N + d + q + 16 <= N + 2 dmax + 16 for the diagonal-sum tiles, below
            // N + d + 15 q + 16 <= N + dmax + 361 for the Toeplitz tiles, q < 24): those
            // samples meet zero template entries
            const int64_t zend =
                min(p.xstride, n + max(2 * (int64_t)dmax + 96, (int64_t)dmax + 416));
            for (int64_t j = np_ + tid; j < zend; j += kBlock) sink.zero(j);
        }
        __syncthreads();
    }
}

// Register-blocked VALU windows of one (period, width) item (strides the matrix-core path
// does not take, and the float64 mode), returning the lane's first minimum.
template <class Real, bool ConstW, class XS>
__device__ __forceinline__ ArgMin valu_width(XS X, const Real* T, const double* __restrict__ D,
                                         WidthArgs a, int32_t R) {
    const int lane = lane_id();
    const int32_t d = a.d, q = a.q;
    const int64_t K = a.K;
    ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};
    for (int64_t base = 0; base < K; base += (int64_t)kWave * R) {
        const int64_t k0 = base + (int64_t)lane * R;
        const int32_t rc = (int32_t)max((int64_t)0, min((int64_t)R, K - k0));
        Real A[kMaxBlock], B[kMaxBlock];
#pragma unroll
        for (int r = 0; r < kMaxBlock; ++r) { A[r] = 0; B[r] = 0; }
        const int64_t s0 = (rc > 0 ? k0 : 0) * q;
        if (q == 1 && R == kMaxBlock)
            window_sums_q1<Real, ConstW>(X + s0, T, d, A, B);
        else
            window_sums_seg<Real, ConstW>(X + s0, T, d, q, R, A, B);
#pragma unroll
        for (int r = 0; r < kMaxBlock; ++r) {
            if (r < rc) {
                const int64_t s = (k0 + r) * q;
                const double m = (D[s + d - 1] - (s ? D[s - 1] : 0.0)) / (double)d;
                if (m > a.depth_min) {
                    const Real rs = (Real)(m * a.os2);
                    double delta;
                    if constexpr (ConstW)
                        delta = a.wc * ((double)(rs * rs) * a.ssk - 2.0 * (double)rs * (double)A[r]);
                    else
                        delta = (double)(rs * rs * B[r] - (Real)2 * rs * A[r]);
                    consider(best, a.chi2f + delta, a.k, (int32_t)s);
                } else {
                    consider(best, a.sentinel, a.k, (int32_t)s);
                }
            }
        }
    }
    return best;
}

// Stage 3 for one (period, width) item, by one wavefront: every candidate position
// s = 0, q, 2q, ... (q = skip stride, GPUFun.cu:397) of width k gets the gate m > depth_min
// on its window mean and the statistic chi2_flat + r^2 B - 2 r A; returns the wave's
// first minimum in (width, position) order.
template <class Real>
__device__ __forceinline__ ArgMin search_width(const SearchParams<Real>& p, int64_t i,
                                               const PeriodMeta& pm, int32_t k) {
    const int lane = lane_id();
    const int c = pm.c;
    const int64_t n = p.pt_off[c + 1] - p.pt_off[c];
    const double wc = p.wconst[c];
    const bool constw = wc == wc;
    const Real* ub = (const Real*)(p.xbuf + i * p.xstride);
    const Planar<Real> vw{ub, ub + p.xstride};
    const double* D = p.dbuf + i * p.xstride;
    ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};

    const int64_t wg = p.wid_off[c] + k;                 // global width index
    const int32_t d = uniform(p.widths[wg]);
    WidthArgs a;
    a.chi2f = p.chi2_flat[c];
    a.sentinel = (double)n;
    a.wc = wc;
    a.depth_min = p.depth_min;
    a.k = k;
    if (d <= 0) {  // empty window: mean 0/0, every position carries the sentinel
        if (lane == 0) consider(best, a.sentinel, k, 0);
        return best;
    }
    const int32_t q = (int32_t)((d > p.skip_point) ? (d / p.skip_point) : 1);
    a.d = d;
    a.q = q;
    a.os2 = 2.0 * p.overshoot[wg];
    a.ssk = p.ss[wg];
    a.inv_d = 1.0 / (double)d;
    a.K = (n + q - 1) / q;
    if (q > 1 && n > 1 && lane == 0) consider(best, a.sentinel, k, 1);
    const int32_t F = d / q;
    ArgMin w;
    if constexpr (sizeof(Real) == 4) {
        // the float16 pair planes of this period (windows.h)
        const char* base = (const char*)(p.xbuf + i * p.xstride);
        const float2 sc = p.xsc[c];
        HPlanes hp;
        hp.X = base;
        hp.W = base + 4 * p.xstride;
        hp.inv_x = 1.f / sc.x;
        hp.inv_w = 1.f / sc.y;
        if (p.use_mfma && q < p.toep_q) {
            const int64_t toff = uniform(p.tmpl_off[wg] + 2 * kTmplPad * wg + kTmplPad);
            const char* T = p.tmplh + 4 * toff;
            if (constw)
                w = toep_width<true>(hp, T, nullptr, D, a);
            else
                w = toep_width<false>(hp, T, T + p.tlo, D, a);
            if (argmin_less(w, best)) best = w;
            return best;
        }
        if (p.use_mfma && q >= 4 && F + 1 <= 16) {
            const int64_t co = uniform(chunk_offset(p.tmpl_off[wg], wg));
            const char* C = p.tmplc + 4 * co;
            const int32_t zrel = (int32_t)(4 * (p.zoff - co));
            // template operands staged in LDS when they fit
            const int32_t ng = chunk_pad(q) / kKGroup;
            if (constw && F == 8 && q >= p.quad_q) {
                if (ng <= kStageGroupsConstW)
                    w = mfma_width<true, true, true, 8, false>(hp, C, nullptr, zrel, D, a);
                else
                    w = mfma_width<true, true, false, 8, false>(hp, C, nullptr, zrel, D, a);
            } else if (constw && F == 8) {
                if (ng <= kStageGroupsConstW)
                    w = mfma_width<true, true, true, 8>(hp, C, nullptr, zrel, D, a);
                else
                    w = mfma_width<true, true, false, 8>(hp, C, nullptr, zrel, D, a);
            } else if (constw && F + 1 < kOnesCol) {
                if (ng <= kStageGroupsConstW)
                    w = mfma_width<true, true, true>(hp, C, nullptr, zrel, D, a);
                else
                    w = mfma_width<true, true, false>(hp, C, nullptr, zrel, D, a);
            } else if (constw) {
                if (ng <= kStageGroupsConstW)
                    w = mfma_width<true, false, true>(hp, C, nullptr, zrel, D, a);
                else
                    w = mfma_width<true, false, false>(hp, C, nullptr, zrel, D, a);
            } else if (q < kQuadQWeighted) {
                // strides this short fit the staged template operands
                static_assert(chunk_pad(kQuadQWeighted - 1) / kKGroup <= kStageGroupsWeighted,
                              "Direct weighted strides must fit the staged template");
                w = mfma_width<false, false, true, 0, true>(hp, C, C + p.clo, zrel, D, a);
            } else if (ng <= kStageGroupsWeighted) {
                w = mfma_width<false, false, true>(hp, C, C + p.clo, zrel, D, a);
            } else {
                w = mfma_width<false, false, false>(hp, C, C + p.clo, zrel, D, a);
            }
            if (argmin_less(w, best)) best = w;
            return best;
        }
    }
    const int64_t toff = uniform(p.tmpl_off[wg] + 2 * kTmplPad * wg + kTmplPad);
    // candidates per lane: enough to cover K with one pass when possible, and (R - 1) q
    // within the buffer's padding of max_width samples
    int32_t R = (int32_t)min((int64_t)kMaxBlock, max((int64_t)1, (a.K + kWave - 1) / kWave));
    R = uniform(min(R, 1 + p.max_width / q));
    if constexpr (sizeof(Real) == 4) {
        const char* base = (const char*)(p.xbuf + i * p.xstride);
        const float2 sc = p.xsc[c];
        const HalfX hx{(const uint32_t*)base, 1.f / sc.x};
        if (constw) {
            w = valu_width<Real, true>(hx, p.tmpl1 + toff, D, a, R);
        } else {
            const HalfX hw{(const uint32_t*)base + p.xstride, 1.f / sc.y};
            w = valu_width<Real, false>(HalfVW{hx, hw}, (const Real*)(p.tmpl2 + toff), D, a, R);
        }
    } else {
        if (constw)
            w = valu_width<Real, true>(ub, p.tmpl1 + toff, D, a, R);
        else
            w = valu_width<Real, false>(vw, (const Real*)(p.tmpl2 + toff), D, a, R);
    }
    if (argmin_less(w, best)) best = w;
    return best;
}

// Stage 3 of the chunk: persistent wavefronts, each taking (period, width slot) items
// from the queue of its XCD, kItemBatch consecutive items per atomic -- the periods
// i = x, x + 8, ... of the chunk go to XCD x, so the widths of one period are searched
// together and its samples stay in that XCD's L2 -- then from the other XCDs' queues once
// its own is drained. Every wavefront leaves after it found all eight queues empty.
template <class Real>
__global__ __launch_bounds__(kBlock, 2) void windows_kernel(SearchParams<Real> p) {
    const int lane = lane_id();
    // One fetch per iteration and no early exits: the loop ends once all eight queues were
    // seen empty (each queue is visited until it is, in order x0, x0 + 1, ...).
    int x = uniform(xcd_id());
    int drained = 0;
    while (drained < 8) {
        unsigned long long it = 0;
        if (lane == 0)
            it = atomicAdd(&p.counters[(1 + x) * kCounterPitch], (unsigned long long)kItemBatch);
        const int64_t item0 = uniform((int64_t)it);
        const int64_t nper = (p.chunk - x + 7) / 8;
        const int64_t nitems = nper > 0 ? nper * p.kmax : 0;
        const int64_t iend = min(item0 + kItemBatch, nitems);
        for (int64_t item = item0; item < iend; ++item) {
            const int64_t i = x + 8 * (item / p.kmax);
            const int32_t j = (int32_t)(item % p.kmax);
            PeriodMeta pm = p.meta[i];
            pm.c = uniform(pm.c);
            pm.status = uniform(pm.status);
            pm.np = uniform(pm.np);
            const int64_t g = p.p0 + i;
            const int32_t k = uniform(p.k_lo[g] + j);
            const int32_t khi = uniform(p.k_hi[g]);
            if (pm.status == 0 && k <= khi) {
                ArgMin best = search_width<Real>(p, i, pm, k);
                best = wave_argmin(best);
                if (lane == 0) p.part[i * p.kmax + j] = best;
            }
        }
        if (item0 >= nitems) {
            ++drained;
            x = (x + 1) & 7;
        }
    }
}

// Stage 4: first minimum over each period's widths in (width, position) order
// (cp.argmin, core.py:866-875); one thread per period of the chunk.
template <class Real>
__global__ void reduce_kernel(SearchParams<Real> p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.chunk) return;
    const int64_t g = p.p0 + i;
    const PeriodMeta pm = p.meta[i];
    if (pm.status != 0) {
        p.chi2[g] = __builtin_nan("");
        p.best_k[g] = pm.status;
        p.best_s[g] = pm.status;
        return;
    }
    ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};
    const int32_t cnt = p.k_hi[g] - p.k_lo[g] + 1;
    for (int32_t j = 0; j < cnt; ++j) {
        const ArgMin c = p.part[i * p.kmax + j];
        if (argmin_less(c, best)) best = c;
    }
    const bool found = best.k != 0x7fffffff;
    p.chi2[g] = found ? best.v : __builtin_nan("");
    p.best_k[g] = found ? best.k : -1;
    p.best_s[g] = found ? best.s : -1;
}

// Stable phase permutations only (bit-exact checks of the sort). Slot layout: phases
// (8 n), merge scratch (8 n + 4 n), bucket lists (8 kMaxBuckets), bucket of each sample
// (2 n).
__host__ __device__ inline int64_t fold_sort_slot_bytes(int64_t n) {
    return ((8 * n + 255) / 256 + (12 * n + 255) / 256 + (8 * (int64_t)kMaxBuckets + 255) / 256) *
           256;
}

__global__ __launch_bounds__(kBlock) void fold_sort_kernel(const double* t, int64_t n,
                                                          const double* periods,
                                                          int64_t n_periods, int neg,
                                                          char* slots, int32_t* perm_out) {
    __shared__ __attribute__((aligned(16))) int hist[kHistLen];
    __shared__ __attribute__((aligned(16))) double s_phi[kWinCap];
    __shared__ int32_t s_idx[kWinCap];
    __shared__ int64_t s_mids[kWinMid];
    __shared__ uint8_t s_wtab[kMaxBuckets];
    __shared__ int s_wb[kMaxWin + 1], s_ws[kMaxWin + 1], s_wcur[kMaxWin];
    __shared__ int red_i[kBlock / kWave];
    __shared__ int s_counts[4];
    char* slot = slots + (int64_t)blockIdx.x * fold_sort_slot_bytes(n);
    double* key = (double*)slot;
    double* tmp_key = (double*)(slot + (8 * n + 255) / 256 * 256);
    int32_t* tmp_id = (int32_t*)(tmp_key + n);
    int64_t* lists = (int64_t*)(slot + (8 * n + 255) / 256 * 256 + (12 * n + 255) / 256 * 256);
    for (int64_t pi = blockIdx.x; pi < n_periods; pi += gridDim.x) {
        int32_t* out = perm_out + pi * n;
        const bool windowed = wg_fold_sort_windows(
            t, n, periods[pi], neg != 0, tmp_id, hist, red_i, s_counts, s_phi, s_idx, s_mids,
            s_wtab, s_wb, s_ws, s_wcur, [&](int wstart, int wcnt) {
                for (int j = threadIdx.x; j < wcnt; j += blockDim.x) out[wstart + j] = s_idx[j];
            });
        if (!windowed)
            wg_fold_sort(t, n, periods[pi], neg != 0, key, out, tmp_key, tmp_id, lists, hist,
                         red_i, s_counts);
        __syncthreads();
    }
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

// Template value sv at pair index i of the two float16 (hi, lo) pair planes from th (plen
// pairs each): S and S^2, S rounded to float32 first.
__device__ __forceinline__ void put_template_pairs(uint32_t* th, int64_t plen, int64_t i,
                                                   double sv) {
    const float s1 = (float)sv;
    th[i] = split_pair(s1);
    th[i + plen] = split_pair((float)(sv * sv));
}

// Search-precision template rows, (S, S^2) pairs, S alone and S^2 alone, each row surrounded by
// kTmplPad zeros: row w starts at tmpl_off[w] + 2 kTmplPad w + kTmplPad. One block per row.
// Float32: also the float16 (hi, lo) pair planes of S and S^2 (S rounded to float32 first)
// for the Toeplitz tiles, two planes of plen pairs from th.
template <class Real>
__global__ void prep_template_kernel(const double* S, const int64_t* off, const int32_t* widths,
                                     int64_t nw, int64_t total_tmpl, int64_t tail,
                                     typename Pair<Real>::type* out, Real* out1, Real* outsq,
                                     uint32_t* th, int64_t plen) {
    auto put = [&](int64_t i, double sv) {
        typename Pair<Real>::type o;
        o.x = (Real)sv;
        o.y = (Real)(sv * sv);
        out[i] = o;
        out1[i] = (Real)sv;
        outsq[i] = (Real)(sv * sv);
        if (th) put_template_pairs(th, plen, i, sv);
    };
    const int64_t w = blockIdx.x;
    if (w >= nw) {
        // the zeroed tail after the last row (over-reads of the window loops)
        const int64_t end = total_tmpl + 2 * kTmplPad * nw;
        for (int64_t i = end + (int64_t)(w - nw) * blockDim.x + threadIdx.x;
             i < end + tail; i += (int64_t)(gridDim.x - nw) * blockDim.x)
            put(i, 0.0);
        return;
    }
    const int32_t d = max(widths[w], 0);
    const int64_t src = off[w];
    const int64_t dst = off[w] + 2 * kTmplPad * w;   // start of the padded row
    for (int32_t i = threadIdx.x; i < d + 2 * kTmplPad; i += blockDim.x) {
        const int32_t j = i - kTmplPad;
        put(dst + i, (j >= 0 && j < d) ? S[src + j] : 0.0);
    }
}

// Chunked template rows of the diagonal-sum tiles (float32) in the two pair planes of
// plen pairs from c (as put_template_pairs), each ending in a zero block of nz pairs at
// zoff; after the F + 1 template chunks, the ones chunks F + 1 (b < q) and F + 2
// (b < d - F q): the pair of 1 in the S plane, 0 in the S^2 plane. One block per width; widths those tiles do
// not take (q < 4 or more than 16 chunks) get no row. The extra blocks past the last width
// clear the zero blocks.
__global__ void chunk_template_kernel(const double* S, const int64_t* off, const int32_t* widths,
                                      int64_t nw, int64_t skip_point, uint32_t* c, int64_t plen,
                                      int64_t zoff, int64_t nz) {
    const int64_t w = blockIdx.x;
    if (w >= nw) {
        for (int64_t i = (int64_t)(w - nw) * blockDim.x + threadIdx.x; i < nz;
             i += (int64_t)(gridDim.x - nw) * blockDim.x)
            put_template_pairs(c, plen, zoff + i, 0.0);
        return;
    }
    const int32_t d = widths[w];
    if (d <= 0) return;
    const int32_t q = (int32_t)((d > skip_point) ? (d / skip_point) : 1);
    const int32_t F = d / q, rem = d - F * q;
    if (q < 4 || F + 1 > 16) return;
    const int32_t qp = chunk_pad(q);
    const int64_t co = chunk_offset(off[w], w);
    const int32_t total = (F + 3) * qp;
    const uint32_t one = split_pair(1.f);
    for (int32_t i = threadIdx.x; i < total; i += blockDim.x) {
        const int32_t ch = i / qp, b = i - ch * qp;
        if (ch <= F) {
            const int32_t len = ch < F ? q : rem;
            put_template_pairs(c, plen, co + i, b < len ? S[off[w] + (int64_t)ch * q + b] : 0.0);
        } else {
            const bool in = b < (ch == F + 1 ? q : rem);
            c[co + i] = in ? one : 0u;
            c[co + i + plen] = 0u;
        }
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

// Flat-model chi2 of each curve, sum (1 - y)^2 / dy^2, float64, the curve's common
// weight 1/dy^2 when every dy is the same value (NaN otherwise), and the power-of-two
// scales of the float32 search's float16 planes (f16_scale of the largest |u| with a
// common weight, else of the largest |u w|, and of the largest w; u and w rounded to
// float32 as the search gathers them); one workgroup per curve.
__global__ __launch_bounds__(kBlock) void chi2_flat_kernel(const int64_t* pt_off,
                                                           const double* y, const double* dy,
                                                           double* out, double* wconst,
                                                           float2* xsc) {
    __shared__ double red[kBlock / kWave];
    __shared__ float red_u[kBlock / kWave], red_v[kBlock / kWave], red_w[kBlock / kWave];
    __shared__ int differ;
    const int c = blockIdx.x;
    const int64_t a = pt_off[c], b = pt_off[c + 1];
    if (threadIdx.x == 0) differ = 0;
    __syncthreads();
    const double e0 = dy[a];
    double s = 0.0;
    float mu = 0.f, mv = 0.f, mw = 0.f;
    int dif = 0;
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) {
        const double u = 1.0 - y[i];
        const double e = dy[i];
        const double w = 1.0 / (e * e);
        s += u * u * w;
        dif |= (e != e0);
        const float uf = (float)u, wf = (float)w;
        mu = fmaxf(mu, fabsf(uf));
        mv = fmaxf(mv, fabsf(uf * wf));
        mw = fmaxf(mw, wf);
    }
    if (dif) atomicOr(&differ, 1);
    s = wave_sum(s);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        mu = fmaxf(mu, __shfl_xor(mu, off, kWave));
        mv = fmaxf(mv, __shfl_xor(mv, off, kWave));
        mw = fmaxf(mw, __shfl_xor(mw, off, kWave));
    }
    if (lane_id() == 0) {
        red[wave_id()] = s;
        red_u[wave_id()] = mu;
        red_v[wave_id()] = mv;
        red_w[wave_id()] = mw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            tot += red[w];
            mu = fmaxf(mu, red_u[w]);
            mv = fmaxf(mv, red_v[w]);
            mw = fmaxf(mw, red_w[w]);
        }
        out[c] = tot;
        wconst[c] = differ ? __builtin_nan("") : 1.0 / (e0 * e0);
        xsc[c] = differ ? make_float2(f16_scale(mv), f16_scale(mw))
                        : make_float2(f16_scale(mu), 1.f);
    }
}

// Per point the float16 (hi, lo) pair of u scaled by its curve's A-sum plane scale, as the
// float32 search's common-weight samples take it; one workgroup per curve (after
// chi2_flat_kernel, which sets the scales).
__global__ __launch_bounds__(kBlock) void pair_points_kernel(const int64_t* pt_off,
                                                             const float2* uw,
                                                             const float2* xsc,
                                                             uint32_t* upair) {
    const int c = blockIdx.x;
    const int64_t a = pt_off[c], b = pt_off[c + 1];
    const float sx = xsc[c].x;
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) upair[i] = split_pair(uw[i].x * sx);
}

}  // namespace gtls
