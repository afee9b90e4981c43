// Smooth limb-darkened transit model for the gradient refinement of grid peaks, with its
// hand-written derivative. It has no counterpart in the reference: the reference stops at
// the grid search and a host-side statistics pass (src/gputls/stats.py:233-488).
//
// Parameters of one model (theta, 6 float64 values): period P, mid-transit time T0,
// duration D (days), depth delta, and Kipping's limb-darkening parameters q1, q2 in (0, 1)
// (u1 = 2 sqrt(q1) q2, u2 = sqrt(q1) (1 - 2 q2)). For a time t:
//     e = t - T0,  n = rint(e / P),  tau = e - n P             (float64: the fold; e / P
//                                                               as e times 1 / P)
//     x = 2 tau / D                                            (chord coordinate)
//     s = 1 - x^2,  W = sigmoid(kappa s),  mu = sqrt(softplus(kappa s) / kappa)
//     I = 1 - u1 (1 - mu) - u2 (1 - mu)^2                      (quadratic law)
//     f = 1 - delta W I
// i.e. a central transit of a small planet across a quadratically limb-darkened star,
// whose disk edge (|x| = 1) is smoothed over ~1/kappa so that f is differentiable in every
// parameter. At the centre f = 1 - delta (to 1e-27); far from transit f = 1. Points with
// kappa s below -kLogCut have W < 1e-26: their deficit and derivatives are set to zero.
//
// Derivatives (n held fixed: it is piecewise constant in P and T0):
//     df/ddelta = -W I,   df/du1 = delta W (1 - mu),   df/du2 = delta W (1 - mu)^2,
//     df/dx = -delta (W' I + W I'),  W' = kappa W (1 - W) (-2 x),
//     I' = (u1 + 2 u2 (1 - mu)) dmu/dx,  dmu/dx = -W x / mu,
//     dx/dT0 = -2 / D,  dx/dP = -2 n / D,  dx/dD = -x / D.
#pragma once
#include <cmath>
#include "wg_primitives.h"

namespace gtls {

constexpr double kTransitKappa = 64.0;  // edge sharpness (1 / smoothing width in s)
constexpr double kLogCut = 60.0;        // kappa s below -kLogCut: outside the model

template <class Real>
struct TransitPoint {
    Real f;                      // model flux
    Real dP, dT0, dD, ddelta, dq1, dq2;  // its partial derivatives
};

// Model flux of one point, and (when Grad) its six partial derivatives. theta as above,
// inv_p = 1 / P; the fold is float64, the rest in Real. The product e inv_p can differ from
// the quotient e / P in the last place, which moves n only for points half a period from
// T0, where the model is 1 with zero derivatives for any duration below P / 2.
template <class Real, bool Grad>
__device__ __forceinline__ TransitPoint<Real> transit_point(double t, const double* theta,
                                                            double inv_p) {
    const double P = theta[0], T0 = theta[1];
    const Real D = (Real)theta[2], delta = (Real)theta[3];
    const Real q1 = (Real)theta[4], q2 = (Real)theta[5];
    const double e = t - T0;
    const double n = rint(e * inv_p);
    const Real tau = (Real)(e - n * P);
    const Real x = (Real)2 * tau / D;
    const Real s = (Real)1 - x * x;
    const Real ks = (Real)kTransitKappa * s;
    TransitPoint<Real> r{(Real)1, 0, 0, 0, 0, 0, 0};
    if (ks < (Real)-kLogCut) return r;
    // W = sigmoid(ks), softplus(ks) = max(ks, 0) + log1p(exp(-|ks|)), both from one exp
    const Real ex = exp(-fabs(ks));
    const Real W = ks >= 0 ? (Real)1 / ((Real)1 + ex) : ex / ((Real)1 + ex);
    const Real sp = (ks > 0 ? ks : (Real)0) + log1p(ex);
    const Real mu = sqrt(sp / (Real)kTransitKappa);
    const Real sq1 = sqrt(q1);
    const Real u1 = (Real)2 * sq1 * q2, u2 = sq1 * ((Real)1 - (Real)2 * q2);
    const Real om = (Real)1 - mu;
    const Real I = (Real)1 - u1 * om - u2 * om * om;
    r.f = (Real)1 - delta * W * I;
    if constexpr (Grad) {
        const Real dWdx = (Real)kTransitKappa * W * ((Real)1 - W) * ((Real)-2 * x);
        const Real dmudx = -W * x / mu;
        const Real dIdx = (u1 + (Real)2 * u2 * om) * dmudx;
        const Real dfdx = -delta * (dWdx * I + W * dIdx);
        const Real two_d = (Real)2 / D;
        r.dT0 = -dfdx * two_d;
        r.dP = -dfdx * two_d * (Real)n;
        r.dD = -dfdx * x / D;
        r.ddelta = -W * I;
        const Real dfdu1 = delta * W * om, dfdu2 = delta * W * om * om;
        // du1/dq1 = q2 / sqrt(q1), du2/dq1 = (1 - 2 q2) / (2 sqrt(q1)),
        // du1/dq2 = 2 sqrt(q1),     du2/dq2 = -2 sqrt(q1)
        r.dq1 = (dfdu1 * q2 + dfdu2 * (Real)0.5 * ((Real)1 - (Real)2 * q2)) / sq1;
        r.dq2 = (dfdu1 - dfdu2) * (Real)2 * sq1;
    }
    return r;
}

// Model m covers the points [pt_off[c], pt_off[c + 1]) of curve c = model_curve[m] and
// writes its fluxes at flux[out_off[m] ..]. Grid: (x blocks over points, y = model).
template <class Real>
__global__ __launch_bounds__(256) void transit_fwd_kernel(const double* __restrict__ theta,
                                                          const int32_t* __restrict__ model_curve,
                                                          const int64_t* __restrict__ pt_off,
                                                          const double* __restrict__ t,
                                                          const int64_t* __restrict__ out_off,
                                                          Real* __restrict__ flux) {
    const int64_t m = blockIdx.y;
    const int c = model_curve[m];
    const int64_t a = pt_off[c], n = pt_off[c + 1] - a;
    double th[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) th[k] = theta[6 * m + k];
    const double inv_p = 1.0 / th[0];
    Real* out = flux + out_off[m];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = transit_point<Real, false>(t[a + i], th, inv_p).f;
}

// grad_theta[m][k] = sum_i grad_flux[out_off[m] + i] * df_i/dtheta_k, one workgroup per
// model; float64 accumulation. Each thread has kBwdUnroll points' loads in flight.
constexpr int kBwdUnroll = 4;
template <class Real>
__global__ __launch_bounds__(256) void transit_bwd_kernel(const double* __restrict__ theta,
                                                          const int32_t* __restrict__ model_curve,
                                                          const int64_t* __restrict__ pt_off,
                                                          const double* __restrict__ t,
                                                          const int64_t* __restrict__ out_off,
                                                          const Real* __restrict__ grad_flux,
                                                          double* __restrict__ grad_theta) {
    __shared__ double red[6][256 / kWave];
    const int64_t m = blockIdx.x;
    const int c = model_curve[m];
    const int64_t a = pt_off[c], n = pt_off[c + 1] - a;
    double th[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) th[k] = theta[6 * m + k];
    const double inv_p = 1.0 / th[0];
    const Real* g = grad_flux + out_off[m];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int64_t nt = blockDim.x;
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += kBwdUnroll * nt) {
        Real gv[kBwdUnroll];
        double tv[kBwdUnroll];
#pragma unroll
        for (int u = 0; u < kBwdUnroll; ++u) {
            const int64_t i = i0 + u * nt;
            const bool in = i < n;
            gv[u] = in ? g[in ? i : 0] : (Real)0;
            tv[u] = in ? t[a + (in ? i : 0)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBwdUnroll; ++u) {
            const Real gi = gv[u];
            if (gi == (Real)0) continue;
            const TransitPoint<Real> p = transit_point<Real, true>(tv[u], th, inv_p);
            acc[0] += (double)(gi * p.dP);
            acc[1] += (double)(gi * p.dT0);
            acc[2] += (double)(gi * p.dD);
            acc[3] += (double)(gi * p.ddelta);
            acc[4] += (double)(gi * p.dq1);
            acc[5] += (double)(gi * p.dq2);
        }
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += red[threadIdx.x][w];
        grad_theta[6 * m + threadIdx.x] = s;
    }
}

// The fit's loss and its gradient in one pass, without flux slots: for model m over the
// points of curve c = model_curve[m] (times t, fluxes y, weights w, all float64),
//     chi2[m] = sum_i w_i (y_i - f_i)^2,
//     grad_theta[6 m + k] = sum_i -2 w_i (y_i - f_i) df_i/dtheta_k   (when Grad).
// One wide workgroup per model (chi2_block threads: the longest curves then leave no tail
// of latency-bound workgroups; 512 for the float64 model, whose derivatives need more
// registers); the residuals and sums are float64 (f in Real). Each thread has kBwdUnroll
// points' loads in flight.
template <class Real>
constexpr int chi2_block() { return sizeof(Real) == 4 ? 1024 : 512; }
template <class Real, bool Grad>
__global__ __launch_bounds__(chi2_block<Real>()) void transit_chi2_kernel(const double* __restrict__ theta,
                                                           const int32_t* __restrict__ model_curve,
                                                           const int64_t* __restrict__ pt_off,
                                                           const double* __restrict__ t,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ w,
                                                           double* __restrict__ chi2,
                                                           double* __restrict__ grad_theta) {
    constexpr int kSums = Grad ? 7 : 1;
    __shared__ double red[kSums][chi2_block<Real>() / kWave];
    const int64_t m = blockIdx.x;
    const int c = model_curve[m];
    const int64_t a = pt_off[c], n = pt_off[c + 1] - a;
    double th[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) th[k] = theta[6 * m + k];
    const double inv_p = 1.0 / th[0];
    double acc[kSums];
#pragma unroll
    for (int k = 0; k < kSums; ++k) acc[k] = 0.0;
    const int64_t nt = blockDim.x;
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += kBwdUnroll * nt) {
        double tv[kBwdUnroll], yv[kBwdUnroll], wv[kBwdUnroll];
#pragma unroll
        for (int u = 0; u < kBwdUnroll; ++u) {
            const int64_t i = i0 + u * nt;
            const bool in = i < n;
            const int64_t j = a + (in ? i : 0);
            tv[u] = in ? t[j] : 0.0;
            yv[u] = in ? y[j] : 1.0;
            wv[u] = in ? w[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBwdUnroll; ++u) {
            const TransitPoint<Real> p = transit_point<Real, Grad>(tv[u], th, inv_p);
            const double r = yv[u] - (double)p.f;
            const double wr = wv[u] * r;
            acc[0] = fma(wr, r, acc[0]);
            if constexpr (Grad) {
                const double g = -2.0 * wr;
                acc[1] = fma(g, (double)p.dP, acc[1]);
                acc[2] = fma(g, (double)p.dT0, acc[2]);
                acc[3] = fma(g, (double)p.dD, acc[3]);
                acc[4] = fma(g, (double)p.ddelta, acc[4]);
                acc[5] = fma(g, (double)p.dq1, acc[5]);
                acc[6] = fma(g, (double)p.dq2, acc[6]);
            }
        }
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < kSums; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < kSums) {
        double s = 0.0;
        for (int v = 0; v < (int)(blockDim.x / kWave); ++v) s += red[threadIdx.x][v];
        if (threadIdx.x == 0) chi2[m] = s;
        else grad_theta[6 * m + threadIdx.x - 1] = s;
    }
}

// The same loss and gradient with the models grouped by curve, in one pass over each
// curve's points. The points of curve c (pt_off) are cut into tiles of kChi2Tile
// consecutive points, tiles [tile_off[c], tile_off[c + 1]) of the launch; the models of
// curve c are [model_off[c], model_off[c + 1]) (at most max_models). Workgroup i takes tile
// i: its kChi2PerThread points per thread stay in registers while every model of the curve
// is evaluated on them. A point whose fold lies more than kNearHalf durations from the
// model's transit (|tau| > kNearHalf D, float64) is outside the model (its kappa s is below
// -kLogCut: flux 1, zero derivatives) and is skipped; its share of chi2, w (y - 1)^2, is
// that of the flat model, which the tile sums once for all models:
//     flat[i] = sum_tile w (y - 1)^2,
//     chi2_m = sum_tiles flat + sum_near w ((y - f)^2 - (y - 1)^2)
//            = sum_tiles flat + sum_near w (1 - f) ((y - f) + (y - 1)).
// The tile's model sums go to
//     part[(i * max_models + j) * S + k]      (model m = model_off[c] + j, k < S),
// S = 7 with Grad (the chi2 difference, then the six gradient components), else 1. Each
// point is read once per call, whatever the number of models. transit_chi2_sum_kernel then
// adds every model's tiles in tile order, so the result does not depend on the launch.
constexpr int kChi2Threads = 256;
constexpr int kChi2PerThread = 8;
constexpr int kChi2Tile = kChi2Threads * kChi2PerThread;
// |x| = 2 |tau| / D above sqrt(1 + kLogCut / kappa) = 1.3919 is outside the model; the
// float64 pre-test keeps a margin over the Real test of transit_point
constexpr double kNearHalf = 0.71;

// Last index c in [0, n) with off[c] <= i, for off ascending with off[0] <= i < off[n].
__device__ __forceinline__ int owner_of(const int64_t* off, int n, int64_t i) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

template <class Real, bool Grad>
__global__ __launch_bounds__(kChi2Threads) void transit_chi2_tile_kernel(
    const double* __restrict__ theta, const int64_t* __restrict__ model_off,
    const int64_t* __restrict__ pt_off, const int64_t* __restrict__ tile_off, int32_t n_curves,
    int32_t max_models, const double* __restrict__ t, const double* __restrict__ y,
    const double* __restrict__ w, double* __restrict__ part, double* __restrict__ flat) {
    constexpr int kSums = Grad ? 7 : 1;
    __shared__ double red[2][kSums][kChi2Threads / kWave];
    __shared__ double red_flat[kChi2Threads / kWave];
    const int64_t i = blockIdx.x;
    const int c = owner_of(tile_off, n_curves, i);
    const int64_t a = pt_off[c] + (i - tile_off[c]) * kChi2Tile;
    const int64_t cnt = min((int64_t)kChi2Tile, pt_off[c + 1] - a);
    const int tid = threadIdx.x;
    // the tile's points (slots past the curve: weight 0, so they add nothing)
    double tv[kChi2PerThread], yv[kChi2PerThread], wv[kChi2PerThread];
#pragma unroll
    for (int u = 0; u < kChi2PerThread; ++u) {
        const int64_t j = (int64_t)u * kChi2Threads + tid;
        const bool in = j < cnt;
        const int64_t g = a + (in ? j : 0);
        tv[u] = in ? t[g] : 0.0;
        yv[u] = in ? y[g] : 1.0;
        wv[u] = in ? w[g] : 0.0;
    }
    const int64_t m0 = model_off[c], m1 = model_off[c + 1];
    const int lane = tid & (kWave - 1), wave = tid / kWave;
    // the flat model's chi2 over the tile (the barrier of the first model orders it)
    {
        double fs = 0.0;
#pragma unroll
        for (int u = 0; u < kChi2PerThread; ++u) {
            const double r0 = yv[u] - 1.0;
            fs = fma(wv[u] * r0, r0, fs);
        }
        fs = wave_sum(fs);
        if (lane == 0) red_flat[wave] = fs;
    }
    for (int64_t m = m0; m < m1; ++m) {
        double th[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) th[k] = theta[6 * m + k];
        const double inv_p = 1.0 / th[0];
        const double near = kNearHalf * th[2];
        double acc[kSums];
#pragma unroll
        for (int k = 0; k < kSums; ++k) acc[k] = 0.0;
        bool any = false;
#pragma unroll
        for (int u = 0; u < kChi2PerThread; ++u) {
            const double e = tv[u] - th[1];
            const double tau = e - rint(e * inv_p) * th[0];
            if (!(fabs(tau) <= near)) continue;   // outside the model: the flat share
            any = true;
            const TransitPoint<Real> pt = transit_point<Real, Grad>(tv[u], th, inv_p);
            const double f = (double)pt.f;
            const double r = yv[u] - f;
            const double wr = wv[u] * r;
            acc[0] = fma(wv[u] * (1.0 - f), r + (yv[u] - 1.0), acc[0]);
            if constexpr (Grad) {
                const double g = -2.0 * wr;
                acc[1] = fma(g, (double)pt.dP, acc[1]);
                acc[2] = fma(g, (double)pt.dT0, acc[2]);
                acc[3] = fma(g, (double)pt.dD, acc[3]);
                acc[4] = fma(g, (double)pt.ddelta, acc[4]);
                acc[5] = fma(g, (double)pt.dq1, acc[5]);
                acc[6] = fma(g, (double)pt.dq2, acc[6]);
            }
        }
        // workgroup sums; the two halves of red alternate between models, so one barrier
        // per model orders every write of red before its reads and after the reads of the
        // model two back. A wavefront with no point near a transit of the model has only
        // zero sums (no term was added) and skips its reductions.
        const int h = (int)((m - m0) & 1);
        if (__ballot(any) != 0ULL) {
#pragma unroll
            for (int k = 0; k < kSums; ++k) {
                const double v = wave_sum(acc[k]);
                if (lane == 0) red[h][k][wave] = v;
            }
        } else if (lane < kSums) {
            red[h][lane][wave] = 0.0;
        }
        __syncthreads();
        if (tid < kSums) {
            double sum = 0.0;
#pragma unroll
            for (int v = 0; v < kChi2Threads / kWave; ++v) sum += red[h][tid][v];
            part[(i * max_models + (m - m0)) * kSums + tid] = sum;
        }
    }
    if (m1 == m0) __syncthreads();
    if (tid == kSums) {
        double sum = 0.0;
#pragma unroll
        for (int v = 0; v < kChi2Threads / kWave; ++v) sum += red_flat[v];
        flat[i] = sum;
    }
}

// chi2[m] (and grad_theta[6 m + k]) as the sum of model m's tile sums in tile order (for
// chi2, then the flat sums of its curve's tiles in tile order); one thread per (model, sum).
template <bool Grad>
__global__ void transit_chi2_sum_kernel(const int64_t* __restrict__ model_off,
                                        const int64_t* __restrict__ tile_off, int32_t n_curves,
                                        int32_t max_models, int64_t n_models,
                                        const double* __restrict__ part,
                                        const double* __restrict__ flat, double* __restrict__ chi2,
                                        double* __restrict__ grad_theta) {
    constexpr int kSums = Grad ? 7 : 1;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_models * kSums) return;
    const int64_t m = idx / kSums;
    const int k = (int)(idx - m * kSums);
    const int c = owner_of(model_off, n_curves, m);
    const int64_t j = m - model_off[c];
    double sum = 0.0;
    for (int64_t i = tile_off[c]; i < tile_off[c + 1]; ++i)
        sum += part[(i * max_models + j) * kSums + k];
    if (k == 0) {
        double fs = 0.0;
        for (int64_t i = tile_off[c]; i < tile_off[c + 1]; ++i) fs += flat[i];
        sum += fs;
    }
    if (k == 0) chi2[m] = sum;
    else grad_theta[6 * m + k - 1] = sum;
}

}  // namespace gtls
