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
// Matrix-core window sums (float32 search). For stride q and width d = F q + rem, the
// window of candidate k covers data blocks k .. k + F (block j = samples jq .. jq + q - 1):
//     G(j, c) = sum_{b<q} S[c q + b] X[j q + b],      A(k) = sum_{c=0}^{F} G(k + c, c),
// (chunk F holds the rem leftover template samples). G is the product of the block matrix
// (rows j, q columns) and the chunk matrix (F + 1 <= 16 columns); it runs on
// v_mfma_f32_16x16x4f32 (operand lanes: A[l % 16][l / 16], B[l / 16][l % 16]; result
// register v of lane l holds C[4 (l / 16) + v][l % 16]).
//
// A super-tile is 4 tiles of 16 G rows (rows 64 T .. 64 T + 63); the 4 tiles share the
// chunk operand. Finished rows go to an 80-row LDS ring per wavefront (64 new rows plus
// the F <= 15 rows of the previous super-tile that its last candidates still need); then
// each lane takes the diagonal sum of one candidate, k = 64 T - F + lane.
//
// The K (b) loop runs in groups of 4 M samples = M MFMA steps; within a group, step s uses
// sample b0 + M (l / 16) + s in K slot l / 16, so each lane's operands for a group are M
// consecutive floats. The loads of the next group are issued before the MFMAs of the
// current one.
// ---------------------------------------------------------------------------------------
typedef float mfma_v4 __attribute__((ext_vector_type(4)));
constexpr int kRingRows = 80;     // 64 rows of a super-tile + up to 15 look-ahead rows
constexpr int kRingStride = 17;   // 16 chunk columns + 1 (bank spread)
constexpr int kRingFloats = kRingRows * kRingStride;

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int M>
struct FV {
    float v[M];
};

// M consecutive floats (one 4 M-byte load).
template <int M>
__device__ __forceinline__ FV<M> ldm(const float* p) {
    FV<M> r;
    __builtin_memcpy(&r, p, sizeof(r));
    return r;
}

// 2 M consecutive floats split into even (first) and odd (second) elements.
template <int M>
__device__ __forceinline__ void ldm2(const float* p, FV<M>& e, FV<M>& o) {
    FV<2 * M> r;
    __builtin_memcpy(&r, p, sizeof(r));
#pragma unroll
    for (int i = 0; i < M; ++i) {
        e.v[i] = r.v[2 * i];
        o.v[i] = r.v[2 * i + 1];
    }
}

template <int M>
__device__ __forceinline__ void maskm(FV<M>& x, int32_t b, int32_t lim) {
#pragma unroll
    for (int i = 0; i < M; ++i) x.v[i] = (b + i < lim) ? x.v[i] : 0.f;
}

__device__ __forceinline__ int ring_wrap(int r) {
    r = r < 0 ? r + kRingRows : r;
    return r >= kRingRows ? r - kRingRows : r;
}

// MFMA steps of stride q with groups of 4 M samples.
__host__ __device__ constexpr int32_t mfma_steps(int32_t q, int32_t m) {
    return m * ((q + 4 * m - 1) / (4 * m));
}

// X: the item's gathered samples (u, or interleaved (v, w) when !ConstW); samples at or
// past np are treated as 0 (read, never used unmasked; reads stay below
// np + d + q + 16). T: the width's template (S, or interleaved (S, S^2)); reads past the
// row (at most q + 15 samples) land in following rows or the zeroed tail of the template
// buffer and are masked. D: prefix sums of u (float64) for the window means.
// emit(k, A, B, Dlo, Dhi) is called by one lane per candidate k < K.
template <bool ConstW, int M, class Emit>
__device__ __forceinline__ void window_sums_mfma_m(const float* __restrict__ X, int64_t np,
                                                   const float* __restrict__ T,
                                                   const double* __restrict__ D, int32_t d,
                                                   int32_t q, int64_t K, float* ringA,
                                                   float* ringB, Emit emit) {
    constexpr int GS = 4 * M;
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    const int32_t F = d / q, rem = d - F * q;
    const int64_t rows = K + F;
    const int64_t nsuper = (rows + 63) / 64;
    const int32_t clen = (col < F) ? q : (col == F ? rem : 0);   // valid b of this chunk
    // columns past the last chunk read chunk 0 (masked)
    const int64_t tcol = (col <= F) ? (int64_t)col * q : 0;
    const int32_t ngroups = (q + GS - 1) / GS;
    int ring0 = 0;   // ring row of G row 64 T
    for (int64_t st = 0; st < nsuper; ++st) {
        // the window-mean prefix sums of this super-tile's candidates, requested early
        const int64_t kc = 64 * st - F + lane;
        const bool cand = kc >= 0 && kc < K;
        double dlo = 0.0, dhi = 0.0;
        if (cand) {
            const int64_t s = kc * q;
            dhi = D[s + d - 1];
            dlo = s ? D[s - 1] : 0.0;
        }
        int64_t abase[4];
        int32_t alim[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t row = 64 * st + 16 * t + col;
            const bool ok = row < rows;
            // rows past the last read block 0 (masked)
            abase[t] = ok ? row * q : 0;
            alim[t] = ok ? (int32_t)min((int64_t)q, max((int64_t)0, np - row * q)) : 0;
        }
        mfma_v4 acc[4], acc2[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            acc[t] = mfma_v4{0.f, 0.f, 0.f, 0.f};
            acc2[t] = mfma_v4{0.f, 0.f, 0.f, 0.f};
        }
        FV<M> a[4], a2[4], bv, bv2;
        auto load_group = [&](int32_t g, FV<M>(&xa)[4], FV<M>(&xa2)[4], FV<M>& xb,
                              FV<M>& xb2) {
            const int32_t b = GS * g + M * kk;
            if constexpr (ConstW) {
                xb = ldm<M>(T + tcol + b);
#pragma unroll
                for (int t = 0; t < 4; ++t) xa[t] = ldm<M>(X + abase[t] + b);
            } else {
                ldm2<M>(T + 2 * (tcol + b), xb, xb2);
#pragma unroll
                for (int t = 0; t < 4; ++t) ldm2<M>(X + 2 * (abase[t] + b), xa[t], xa2[t]);
            }
            maskm<M>(xb, b, clen);
            if constexpr (!ConstW) maskm<M>(xb2, b, clen);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                maskm<M>(xa[t], b, alim[t]);
                if constexpr (!ConstW) maskm<M>(xa2[t], b, alim[t]);
            }
        };
        load_group(0, a, a2, bv, bv2);
        for (int32_t g = 0; g < ngroups; ++g) {
            FV<M> na[4], na2[4], nb, nb2;
            const bool more = g + 1 < ngroups;
            if (more) load_group(g + 1, na, na2, nb, nb2);
#pragma unroll
            for (int i = 0; i < M; ++i) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t].v[i], bv.v[i], acc[t],
                                                                  0, 0, 0);
                    if constexpr (!ConstW)
                        acc2[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[t].v[i], bv2.v[i],
                                                                       acc2[t], 0, 0, 0);
                }
            }
            if (more) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    a[t] = na[t];
                    if constexpr (!ConstW) a2[t] = na2[t];
                }
                bv = nb;
                if constexpr (!ConstW) bv2 = nb2;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int r = ring0 + 16 * t + 4 * kk + v;
                r = r >= kRingRows ? r - kRingRows : r;
                ringA[r * kRingStride + col] = acc[t][v];
                if constexpr (!ConstW) ringB[r * kRingStride + col] = acc2[t][v];
            }
        }
        wave_sync_lds();
        if (cand) {
            float sa = 0.f, sb = 0.f;
            const int r0 = ring0 - F + lane;
            for (int32_t c = 0; c <= F; ++c) {
                const int r = ring_wrap(r0 + c);
                sa += ringA[r * kRingStride + c];
                if constexpr (!ConstW) sb += ringB[r * kRingStride + c];
            }
            emit(kc, sa, sb, dlo, dhi);
        }
        wave_sync_lds();
        ring0 = ring0 + 64 >= kRingRows ? ring0 + 64 - kRingRows : ring0 + 64;
    }
}

// Picks the group size (4, 8 or 16 samples) with the fewest MFMA steps for stride q,
// preferring larger groups (fewer, wider loads) on ties.
template <bool ConstW, class Emit>
__device__ __forceinline__ void window_sums_mfma(const float* __restrict__ X, int64_t np,
                                                 const float* __restrict__ T,
                                                 const double* __restrict__ D, int32_t d,
                                                 int32_t q, int64_t K, float* ringA,
                                                 float* ringB, Emit emit) {
    const int32_t s1 = mfma_steps(q, 1), s2 = mfma_steps(q, 2), s4 = mfma_steps(q, 4);
    if (s4 <= s1 && s4 <= s2)
        window_sums_mfma_m<ConstW, 4>(X, np, T, D, d, q, K, ringA, ringB, emit);
    else if (s2 <= s1)
        window_sums_mfma_m<ConstW, 2>(X, np, T, D, d, q, K, ringA, ringB, emit);
    else
        window_sums_mfma_m<ConstW, 1>(X, np, T, D, d, q, K, ringA, ringB, emit);
}

}  // namespace gtls
