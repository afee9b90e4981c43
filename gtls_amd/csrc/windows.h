// In-transit part of the TLS statistic for every candidate position of one width.
//
// Candidates are the positions s = k q (q = skip stride) of the phase-sorted sequence.
// For candidate s the window sums are
//     A(s) = sum_{i<d} S_i X_{s+i},   B(s) = sum_{i<d} S2_i W_{s+i}
// (X = u when every sample has the same weight, X = v = u w otherwise). One wavefront
// owns one width; each lane owns R <= 8 consecutive candidates s_r = s_0 + r q, whose
// windows overlap by d - q samples, so one load of X_j feeds up to R accumulators. The
// lane walks its data in segments of q samples: in segment c, candidate r uses template
// indices (c - r) q + b, b < q, which are the same for every lane of the wavefront --
// the template values are scalar loads and the only per-lane memory stream is the lane's
// own contiguous data range. d, q, R and the template pointer must be wave-uniform.
//
// Templates are stored with kTmplPad zero samples on both sides of each row, so the
// stride-1 path can read a fixed 15-sample span without bounds tests.
#pragma once
#include "wg_primitives.h"

namespace gtls {

constexpr int kMaxBlock = 8;   // candidates per lane
constexpr int kTmplPad = 16;   // zero samples before and after each template row

// Template rows are read through the constant address space: they are never written by
// the kernel, and uniform loads from it compile to scalar (SMEM) loads.
template <class T>
using const_ptr = const T __attribute__((address_space(4)))*;

template <class T>
__device__ __forceinline__ const_ptr<T> as_const(const T* p) {
    return (const_ptr<T>)p;
}

// One template entry: S (and S^2 when the weights differ). Template rows are Real arrays,
// interleaved (S, S^2) for the weighted path.
template <class Real>
struct TV {
    Real s, s2;
};

template <bool ConstW, class Real>
__device__ __forceinline__ TV<Real> tload(const_ptr<Real> T, int32_t i) {
    if constexpr (ConstW) return TV<Real>{T[i], (Real)0};
    else return TV<Real>{T[2 * i], T[2 * i + 1]};
}

template <bool ConstW, class Real, class XT>
__device__ __forceinline__ void acc1(Real& A, Real& B, const TV<Real> t, const XT x) {
    if constexpr (ConstW) {
        A = fma(t.s, x, A);
    } else {
        A = fma(t.s, x.x, A);
        B = fma(t.s2, x.y, B);
    }
}

// Stride q >= 2 (or any q with fewer than 8 candidates per lane): segment walk. Per group
// of four samples the template runs of four candidates are fetched together (indices
// clamped into the padded row, so fetches for inactive candidates stay in bounds and are
// simply not used), then the active candidates accumulate.
template <class Real, bool ConstW, class XT>
__device__ __forceinline__ void window_sums_seg(const XT* __restrict__ X, const Real* Tg,
                                                int32_t d, int32_t q, int32_t R, Real* A, Real* B) {
    const const_ptr<Real> T = as_const(Tg);
    const int32_t F = d / q, rem = d - F * q;
    const int32_t nseg = R + F;
    const int32_t q4 = q & ~3;
    const int32_t ilo = -kTmplPad, ihi = d + kTmplPad - 4;
    for (int32_t c = 0; c < nseg; ++c) {
        const int32_t rlo = max(0, c - F + 1), rhi = min(R - 1, c);
        const XT* xs = X + (int64_t)c * q;
        if (rlo <= rhi) {
            int32_t b = 0;
            for (; b < q4; b += 4) {
                const XT x0 = xs[b], x1 = xs[b + 1], x2 = xs[b + 2], x3 = xs[b + 3];
#pragma unroll
                for (int rg = 0; rg < kMaxBlock; rg += 4) {
                    if (rg > rhi || rg + 3 < rlo) continue;
                    TV<Real> t[4][4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int32_t i = min(max((c - rg - r) * q + b, ilo), ihi);
#pragma unroll
                        for (int e = 0; e < 4; ++e) t[r][e] = tload<ConstW>(T, i + e);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (rg + r >= rlo && rg + r <= rhi) {
                            acc1<ConstW>(A[rg + r], B[rg + r], t[r][0], x0);
                            acc1<ConstW>(A[rg + r], B[rg + r], t[r][1], x1);
                            acc1<ConstW>(A[rg + r], B[rg + r], t[r][2], x2);
                            acc1<ConstW>(A[rg + r], B[rg + r], t[r][3], x3);
                        }
                    }
                }
            }
            for (; b < q; ++b) {
                const XT x = xs[b];
#pragma unroll
                for (int r = 0; r < kMaxBlock; ++r)
                    if (r >= rlo && r <= rhi)
                        acc1<ConstW>(A[r], B[r], tload<ConstW>(T, (c - r) * q + b), x);
            }
        }
        // the partial last chunk of candidate r = c - F: template indices F q + b, b < rem
        const int32_t re = c - F;
        if (rem > 0 && re >= 0 && re < R) {
            for (int32_t b = 0; b < rem; ++b) {
                const XT x = xs[b];
                const TV<Real> t = tload<ConstW>(T, F * q + b);
#pragma unroll
                for (int r = 0; r < kMaxBlock; ++r)
                    if (r == re) acc1<ConstW>(A[r], B[r], t, x);
            }
        }
    }
}

// Stride 1 with 8 candidates per lane: sample j feeds candidates r with template index
// j - r; eight samples at a time use the zero-padded template span [j0 - 7, j0 + 7].
template <class Real, bool ConstW, class XT>
__device__ __forceinline__ void window_sums_q1(const XT* __restrict__ X, const Real* Tg,
                                               int32_t d, Real* A, Real* B) {
    const const_ptr<Real> T = as_const(Tg);
    const int32_t L = d + kMaxBlock - 1;
    for (int32_t j0 = 0; j0 < L; j0 += 8) {
        XT x[8];
        TV<Real> t[15];
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = X[j0 + e];
#pragma unroll
        for (int e = 0; e < 15; ++e) t[e] = tload<ConstW>(T, j0 - 7 + e);
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int r = 0; r < kMaxBlock; ++r) acc1<ConstW>(A[r], B[r], t[e + 7 - r], x[e]);
    }
}

}  // namespace gtls
