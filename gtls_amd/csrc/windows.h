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
            // the next group of four samples is requested before the current one is used
            XT n0, n1, n2, n3;
            if (q4 > 0) { n0 = xs[0]; n1 = xs[1]; n2 = xs[2]; n3 = xs[3]; }
            for (; b < q4; b += 4) {
                const XT x0 = n0, x1 = n1, x2 = n2, x3 = n3;
                if (b + 4 < q4) {
                    n0 = xs[b + 4]; n1 = xs[b + 5]; n2 = xs[b + 6]; n3 = xs[b + 7];
                }
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

namespace gtls {

// ---------------------------------------------------------------------------------------
// Matrix-core window sums (float32 search): with blocks of q samples and template chunks
// of q samples,
//     G(j, c) = sum_{b<q} S[c q + b] X[j q + b],      A(k) = sum_{c=0}^{F} G(k + c, c),
// (chunk F holds the d - F q leftover template samples). G is the product of the block
// matrix U (rows j, q columns) and the chunk matrix (F + 1 <= 16 columns): it runs on
// v_mfma_f32_16x16x4f32, 16 blocks x 16 chunks per tile, the K loop walking b in steps
// of 4. Operand lanes: A[l % 16][l / 16], B[l / 16][l % 16]; result register v of lane l
// holds C[4 (l / 16) + v][l % 16]. Finished tiles go to a 32-row LDS ring per wavefront;
// after tile t the 16 candidates whose last block falls in it take their diagonal sums.
// ---------------------------------------------------------------------------------------
typedef float mfma_v4 __attribute__((ext_vector_type(4)));
constexpr int kRingRows = 32;
constexpr int kRingStride = 17;   // 16 chunk columns + 1 (bank spread)

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads four consecutive floats (16 B) of a sample or template stream.
__device__ __forceinline__ float4 ld4(const float* p) {
    float4 v;
    __builtin_memcpy(&v, p, sizeof(v));
    return v;
}

__device__ __forceinline__ float4 mask4(float4 v, int32_t b, int32_t lim) {
    v.x = (b + 0 < lim) ? v.x : 0.f;
    v.y = (b + 1 < lim) ? v.y : 0.f;
    v.z = (b + 2 < lim) ? v.z : 0.f;
    v.w = (b + 3 < lim) ? v.w : 0.f;
    return v;
}

// X: the item's gathered samples (u, or interleaved (v, w) when !ConstW); samples at or
// past np are treated as 0 (they are read but never used unmasked).
// T: the width's template (S, or interleaved (S, S^2)); reads past the row land in
// following rows or the zeroed tail of the template buffer and are masked.
// D: prefix sums of u (float64), used for the candidates' window means.
// emit(k, A, B, Dlo, Dhi) is called by one lane per candidate k < K.
//
// The K (b) loop runs in groups of 16 samples = 4 MFMA steps; within a group, step s uses
// sample b0 + 4 (l / 16) + s in K slot l / 16, so each lane's operands for the group are 4
// consecutive floats (one 16-byte load per stream). The loads of the next group are issued
// before the MFMAs of the current one.
template <bool ConstW, class Emit>
__device__ __forceinline__ void window_sums_mfma(const float* __restrict__ X, int64_t np,
                                                 const float* __restrict__ T,
                                                 const double* __restrict__ D, int32_t d,
                                                 int32_t q, int64_t K, float* ringA,
                                                 float* ringB, Emit emit) {
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    const int32_t F = d / q, rem = d - F * q;
    const int64_t rows = K + F;
    const int64_t ntiles = (rows + 15) / 16;
    const int32_t clen = (col < F) ? q : (col == F ? rem : 0);   // valid b of this chunk
    // columns past the last chunk read row 0 (masked): keeps reads within d + q + 16
    const int64_t tcol = (col <= F) ? (int64_t)col * q : 0;
    const int32_t ngroups = (q + 15) / 16;
    for (int64_t t = 0; t < ntiles; ++t) {
        // the window-mean prefix sums of this tile's candidates, requested early
        const int64_t kc = 16 * t - F + lane;
        const bool cand = lane < 16 && kc >= 0 && kc < K;
        double dlo = 0.0, dhi = 0.0;
        if (cand) {
            const int64_t s = kc * q;
            dhi = D[s + d - 1];
            dlo = s ? D[s - 1] : 0.0;
        }
        const int64_t row = 16 * t + col;
        const bool rowok = row < rows;
        // rows past the last read block 0 (masked): keeps reads below n + d + q + 16
        const int64_t abase = rowok ? row * q : 0;
        mfma_v4 acc = {0.f, 0.f, 0.f, 0.f};
        mfma_v4 acc2 = {0.f, 0.f, 0.f, 0.f};
        // group g covers samples [16 g, 16 g + 16) of every block; lane part: 4 (l/16)
        auto load_group = [&](int32_t g, float4& a, float4& a2, float4& bv, float4& bv2) {
            const int32_t b = 16 * g + 4 * kk;
            const int64_t ai = abase + b;
            if constexpr (ConstW) {
                a = ld4(X + ai);
                bv = ld4(T + tcol + b);
            } else {
                const float4 p0 = ld4(X + 2 * ai), p1 = ld4(X + 2 * ai + 4);
                a = make_float4(p0.x, p0.z, p1.x, p1.z);
                a2 = make_float4(p0.y, p0.w, p1.y, p1.w);
                const float4 q0 = ld4(T + 2 * (tcol + b)), q1 = ld4(T + 2 * (tcol + b) + 4);
                bv = make_float4(q0.x, q0.z, q1.x, q1.z);
                bv2 = make_float4(q0.y, q0.w, q1.y, q1.w);
            }
            // A: samples past the block (b >= q) or past the data are zero; B: past the
            // chunk's length
            const int32_t alim = rowok ? (int32_t)min((int64_t)q, max((int64_t)0, np - abase)) : 0;
            a = mask4(a, b, alim);
            bv = mask4(bv, b, clen);
            if constexpr (!ConstW) {
                a2 = mask4(a2, b, alim);
                bv2 = mask4(bv2, b, clen);
            }
        };
        float4 a, a2, bv, bv2;
        load_group(0, a, a2, bv, bv2);
        for (int32_t g = 0; g < ngroups; ++g) {
            float4 na, na2, nb, nb2;
            if (g + 1 < ngroups) load_group(g + 1, na, na2, nb, nb2);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, bv.x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, bv.y, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, bv.z, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, bv.w, acc, 0, 0, 0);
            if constexpr (!ConstW) {
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2.x, bv2.x, acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2.y, bv2.y, acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2.z, bv2.z, acc2, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2.w, bv2.w, acc2, 0, 0, 0);
            }
            if (g + 1 < ngroups) {
                a = na;
                bv = nb;
                if constexpr (!ConstW) {
                    a2 = na2;
                    bv2 = nb2;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = (int)((16 * t + 4 * kk + v) & (kRingRows - 1));
            ringA[r * kRingStride + col] = acc[v];
            if constexpr (!ConstW) ringB[r * kRingStride + col] = acc2[v];
        }
        wave_sync_lds();
        if (cand) {
            float sa = 0.f, sb = 0.f;
            for (int32_t c = 0; c <= F; ++c) {
                const int r = (int)((kc + c) & (kRingRows - 1));
                sa += ringA[r * kRingStride + c];
                if constexpr (!ConstW) sb += ringB[r * kRingStride + c];
            }
            emit(kc, sa, sb, dlo, dhi);
        }
        wave_sync_lds();
    }
}

}  // namespace gtls
