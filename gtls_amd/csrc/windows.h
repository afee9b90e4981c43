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
#include <type_traits>
#include "wg_primitives.h"

namespace gtls {

template <class Real> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

constexpr int kMaxBlock = 8;   // candidates per lane

// Wave-uniform values of one (period, width) item that the window routines need to turn
// window sums into the statistic: chi2_flat, the sentinel N, the common weight, sum S^2,
// 2 overshoot, 1 / d, the depth gate; the number of candidates K, the width d, its
// stride q and its index k.
struct WidthArgs {
    double chi2f, sentinel, wc, ssk, os2, inv_d, depth_min;
    int64_t K;
    int32_t d, q, k;
};

__device__ __forceinline__ void consider(ArgMin& b, double v, int32_t k, int32_t s) {
    ArgMin c{v, k, s};
    if (argmin_less(c, b)) b = c;
}
constexpr int kTmplPad = 384;  // zero samples before and after each template row

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

// Samples of the weighted path in two planes (v = u w and w), read as (v, w) pairs.
template <class Real>
struct Planar {
    const Real* v;
    const Real* w;
    __device__ __forceinline__ typename Pair<Real>::type operator[](int64_t i) const {
        typename Pair<Real>::type r;
        r.x = v[i];
        r.y = w[i];
        return r;
    }
    __device__ __forceinline__ Planar operator+(int64_t o) const { return Planar{v + o, w + o}; }
};

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
template <class Real, bool ConstW, class XS>
__device__ __forceinline__ void window_sums_seg(XS X, const Real* Tg, int32_t d, int32_t q,
                                                int32_t R, Real* A, Real* B) {
    using XT = std::remove_cv_t<std::remove_reference_t<decltype(X[0])>>;
    const const_ptr<Real> T = as_const(Tg);
    const int32_t F = d / q, rem = d - F * q;
    const int32_t nseg = R + F;
    const int32_t q4 = q & ~3;
    const int32_t ilo = -kTmplPad, ihi = d + kTmplPad - 4;
    for (int32_t c = 0; c < nseg; ++c) {
        const int32_t rlo = max(0, c - F + 1), rhi = min(R - 1, c);
        const XS xs = X + (int64_t)c * q;
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
template <class Real, bool ConstW, class XS>
__device__ __forceinline__ void window_sums_q1(XS X, const Real* Tg, int32_t d, Real* A,
                                               Real* B) {
    using XT = std::remove_cv_t<std::remove_reference_t<decltype(X[0])>>;
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
// Float32 search operands in float16 high/low pairs. A value x (scaled by a power of two
// so that |x| < 2^14) is stored as the pair (hi, lo) = (f16(x), f16(x - hi)) in one
// 32-bit word; hi + lo holds x to about 2^-22 relative. Sample and template planes are
// arrays of such pairs. From a template pair (hi(s), lo(s)) the registers form the two
// operands (hi(s), hi(s)) and (lo(s), 0) (one byte permute and one shift), so that with
// the K dimension running over (sample, half) the two products
//     sum_i hi(x_i) hi(s_i) + lo(x_i) hi(s_i)   and   sum_i hi(x_i) lo(s_i)
// are one v_mfma_f32_16x16x32_f16 each per 16 samples (the dropped lo lo term is ~2^-22
// of the product). Every operand is 4 consecutive pairs (16 bytes) starting at a pair
// boundary, so every load is dword-aligned, and the window sums keep float32 accuracy at
// 2 MFMAs of 16 x 16 x 32 per 16 samples instead of 4 of 16 x 16 x 4.
// ---------------------------------------------------------------------------------------
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t f16_pair(_Float16 a, _Float16 b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, a) |
           ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
// The (hi, lo) pair of x.
__device__ __forceinline__ uint32_t split_pair(float x) {
    const _Float16 hi = (_Float16)x;
    return f16_pair(hi, (_Float16)(x - (float)hi));
}
__device__ __forceinline__ float pair_value(uint32_t p) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(p & 0xffffu)) +
           (float)__builtin_bit_cast(_Float16, (uint16_t)(p >> 16));
}

// Power-of-two scale s with max_abs * s < 2^14 (1 for max_abs = 0 or not finite).
__host__ __device__ inline float f16_scale(float max_abs) {
    if (!(max_abs > 0.f) || !(max_abs < 3.0e38f)) return 1.f;
    int e = 0;
    (void)frexpf(max_abs, &e);        // max_abs < 2^e
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    return ldexpf(1.f, 14 - e);
}

// Samples of the float32 VALU windows read from a pair plane: (hi + lo) / scale.
struct HalfX {
    const uint32_t* p;
    float inv;       // 1 / scale
    __device__ __forceinline__ float operator[](int64_t i) const { return pair_value(p[i]) * inv; }
    __device__ __forceinline__ HalfX operator+(int64_t o) const { return HalfX{p + o, inv}; }
};
// ... and the (v, w) pairs of the weighted path.
struct HalfVW {
    HalfX v, w;
    __device__ __forceinline__ float2 operator[](int64_t i) const {
        return make_float2(v[i], w[i]);
    }
    __device__ __forceinline__ HalfVW operator+(int64_t o) const { return HalfVW{v + o, w + o}; }
};


// The running first minimum of one lane over its candidates of one width in the matrix-core
// paths. A lane visits its candidates in increasing position order, so a later candidate
// replaces the minimum only when its statistic is strictly smaller. thr bounds from above
// (statistic - chi2_flat) of the minimum: a candidate whose float estimate of that
// difference, less the estimate's error bound, is not below thr cannot replace it, and its
// exact float64 statistic is skipped (NaN estimates are evaluated exactly). The result is
// the same as evaluating every candidate exactly.
struct LaneMin {
    ArgMin best;
    float thr;
    bool sent_live;   // the sentinel N is below the minimum: a closed gate replaces it
};

__device__ __forceinline__ LaneMin lane_min_init() {
    return LaneMin{ArgMin{__builtin_inf(), 0x7fffffff, 0x7fffffff}, __builtin_inff(), true};
}

__device__ __forceinline__ void lane_take(LaneMin& L, double v, int32_t s, const WidthArgs& a) {
    L.best = ArgMin{v, a.k, s};
    const float dv = (float)(v - a.chi2f);
    L.thr = dv + 1e-6f * fabsf(dv) + 1e-12f * (float)fabs(a.chi2f) + 1e-30f;
    L.sent_live = a.sentinel < v;
}

// Float copies of one width's constants for the estimates (wss = w0 sum S^2, w2 = 2 w0).
struct WidthConstsF {
    float wss, w2, os2, inv_d, depth_min;
};
__device__ __forceinline__ WidthConstsF width_consts_f(const WidthArgs& a) {
    return WidthConstsF{(float)(a.wc * a.ssk), (float)(2.0 * a.wc), (float)a.os2,
                        (float)a.inv_d, (float)a.depth_min};
}

// One candidate at position s with an open (gate) or closed window-mean gate, r = 2 m
// overshoot (rs, float), and window sums sa (A) and, with per-sample weights, sb (B):
// statistic chi2_flat + w0 (r^2 sum S^2 - 2 r A) (common weight w0, float64) or
// chi2_flat + (r^2 B - 2 r A) (float), else the sentinel N.
template <bool ConstW>
__device__ __forceinline__ void tile_candidate(LaneMin& L, const WidthArgs& a,
                                               const WidthConstsF& f, bool gate, float rs,
                                               float sa, float sb, int32_t s) {
    if (gate) {
        if constexpr (ConstW) {
            const float t1 = f.wss * (rs * rs);
            const float t2 = f.w2 * rs * sa;
            const float lo = (t1 - t2) - 1e-6f * (fabsf(t1) + fabsf(t2));
            if (!(lo >= L.thr)) {
                const double v =
                    a.chi2f + a.wc * ((double)(rs * rs) * a.ssk - 2.0 * (double)rs * (double)sa);
                if (v < L.best.v) lane_take(L, v, s, a);
            }
        } else {
            const float df = rs * rs * sb - 2.f * rs * sa;
            if (!(df >= L.thr)) {
                const double v = a.chi2f + (double)df;
                if (v < L.best.v) lane_take(L, v, s, a);
            }
        }
    } else if (L.sent_live) {
        lane_take(L, a.sentinel, s, a);
    }
}

// ---------------------------------------------------------------------------------------
// Matrix-core window sums (float32 search). For stride q and width d = F q + rem, the
// window of candidate k covers data blocks k .. k + F (block j = samples jq .. jq + q - 1):
//     G(j, c) = sum_{b<q} S[c q + b] X[j q + b],      A(k) = sum_{c=0}^{F} G(k + c, c),
// (chunk F holds the rem leftover template samples). G is the product of the block matrix
// (rows j, q columns) and the chunk matrix (F + 1 <= 16 columns); it runs on
// v_mfma_f32_16x16x32_f16 over the pair planes (operand lanes: A[l % 16][8 (l / 16) + e],
// B[8 (l / 16) + e][l % 16], e < 8, i.e. the 4 pairs of samples 4 (l / 16) .. + 3 of a
// K group of 16 samples; result register v of lane l holds C[4 (l / 16) + v][l % 16]).
//
// A super-tile is 4 tiles of 16 G rows (rows 64 T .. 64 T + 63); the 4 tiles share the
// chunk operand. Finished rows go to an 80-row LDS ring per wavefront (64 new rows plus
// the F <= 15 rows of the previous super-tile that its last candidates still need); then
// each lane takes the diagonal sum of one candidate, k = 64 T - F + lane.
//
// The loads of the next K group are issued before the MFMAs of the current one, and the
// first group of the next sweep (the B plane, or the next super-tile) before the last
// MFMAs of this one, so its latency passes during the diagonal sums.
// ---------------------------------------------------------------------------------------
typedef float mfma_v4 __attribute__((ext_vector_type(4)));
constexpr int kRingRows = 80;     // 64 rows of a super-tile + up to 15 look-ahead rows
// Column-major: column c at c * kRingPitch, rows of 4 bytes; rows 0 .. 15 are stored again at
// kRingRows .. kRingRows + 15, so the up to 16 consecutive rows a lane's diagonal reads from
// its first row never wrap (one base address, immediate offsets). 100 = 36 mod 64 banks:
// the 16 columns of a 16-byte row store fall on distinct banks.
constexpr int kRingMirror = 16;
constexpr int kRingPitch = 100;
static_assert(kRingPitch >= kRingRows + kRingMirror, "ring column too short for the mirror");
constexpr int kRingFloats = 16 * kRingPitch;
constexpr int kKGroup = 16;       // samples per K group (one 16x16x32 MFMA step over pairs)

// 4 consecutive pairs (one 16-byte load).
__device__ __forceinline__ h8 ld_h8(const char* p) {
    h8 r;
    __builtin_memcpy(&r, p, sizeof(r));
    return r;
}

// LDS bytes (address space 3: ds_read / ds_write).
typedef __attribute__((address_space(3))) char lds_char;

__device__ __forceinline__ int ring_wrap(int r) {
    r = r < 0 ? r + kRingRows : r;
    return r >= kRingRows ? r - kRingRows : r;
}

// Chunked template rows of the matrix-core path: for width w with stride q, F = d / q and
// q_pad = q rounded up to whole K groups, chunk c <= F sits at chunk_offset(w) + c q_pad
// (in pairs of each plane), holds S[c q + b] for b < its length (q, or d - F q for c = F)
// and zeros up to q_pad; in the S plane, chunks F + 1 and F + 2 hold ones for b < q and
// b < d - F q (the window sums of the samples themselves, in tile columns 14 and 15 when
// F <= 13). Other columns read the zero block that follows the rows of each plane.
// (F + 3) q_pad <= 4 d + 270, so the offsets below leave room for every row.
constexpr int kOnesCol = 14;     // tile column of the ones chunk (F + 1); the rem ones in 15
__host__ __device__ inline int64_t chunk_offset(int64_t tmpl_off, int64_t w) {
    return 4 * tmpl_off + 1024 * w;
}
__host__ __device__ constexpr int32_t chunk_pad(int32_t q) {
    return (q + kKGroup - 1) / kKGroup * kKGroup;
}

// Operand byte offsets of one sweep over NT tiles: the sample pair plane X and the
// template pair plane C, the lane's A offsets (one per tile) and B offset; or, for B
// operands staged in LDS (h_load with BL), the lane's B operand of K group 0 at LB, group
// g 1024 g bytes further.
template <int NT>
struct HSweep {
    const char* X;       // sample pair plane (wave-uniform)
    const char* C;       // template pair plane (wave-uniform)
    uint32_t xo[NT];     // lane's A operand byte offsets, one per tile
    int32_t co;          // lane's B operand byte offset
};

// One K group's operands of a sweep: the sample pairs of the NT tiles and the template
// pairs (expanded into the two MFMA operands by h_mma).
// With Ones, also a third template operand: (1, 1) pairs where the lane's taps fall inside
// the window of its column (a row of d ones shifted like the template: the window sums of
// the samples, Toeplitz tiles), formed in registers from the lane's template tap
// w.co / 4 + 16 g and the width d.
template <int NT, bool Ones = false>
struct HSet {
    h8 a[NT];
    h8 b;
    h8 b3[Ones ? 1 : 0];
};

template <int NT, bool Ones>
__device__ __forceinline__ void h_load(const HSweep<NT>& w, int32_t g, HSet<NT, Ones>& o,
                                       int32_t d = 0) {
    const int32_t off = 4 * kKGroup * g;
#pragma unroll
    for (int t = 0; t < NT; ++t) o.a[t] = ld_h8(w.X + (w.xo[t] + (uint32_t)off));
    o.b = ld_h8(w.C + (int64_t)(w.co + off));
    if constexpr (Ones) {
        const int32_t tap = (w.co + off) / 4;   // exact: w.co and off are multiples of 4
        uint32_t v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            v[e] = ((uint32_t)(tap + e) < (uint32_t)d) ? 0x3C003C00u : 0u;   // (1, 1)
        __builtin_memcpy(&o.b3[0], v, sizeof(v));
    }
}

// The sample operands are loaded by lanes l = 4 r + s (row r, K slice s: each 4-lane quad
// reads one row's 64 contiguous bytes, which the L1 cache serves at up to 4x the rate of
// the operand layout, where the 4 lanes of a quad read 4 different rows) and moved in
// place to the MFMA operand layout (row l % 16, slice l / 16) right before their MFMAs:
// lane l takes the 4 dwords of lane 4 (l % 16) + l / 16.
__device__ __forceinline__ int a_permute_addr() {
    const int lane = lane_id();
    return 4 * (4 * (lane & 15) + (lane >> 4));
}
template <int NT, bool Ones>
__device__ __forceinline__ void a_permute(HSet<NT, Ones>& o, int addr) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        uint32_t v[4];
        __builtin_memcpy(v, &o.a[t], sizeof(v));
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)v[e]);
        __builtin_memcpy(&o.a[t], v, sizeof(v));
    }
}

// K groups of a sweep over `len` taps (taps past `len` meet zero template entries).
__device__ __forceinline__ int32_t k_groups(int32_t len) {
    return (len + kKGroup - 1) / kKGroup;
}

// The two pair products of one K group into the NT tiles' accumulators (each pass over
// the tiles before the next, so that consecutive MFMAs are independent); with Ones, the
// sample sums into accw.
template <int NT, bool Ones>
__device__ __forceinline__ void h_mma(mfma_v4 (&acc)[NT], mfma_v4 (&accw)[NT],
                                      const HSet<NT, Ones>& o) {
    // template operands (hi, hi) and (lo, 0) from the (hi, lo) pairs
    uint32_t p[4], v1[4], v2[4];
    __builtin_memcpy(p, &o.b, sizeof(p));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        v1[e] = __builtin_amdgcn_perm(p[e], p[e], 0x01000100u);
        v2[e] = p[e] >> 16;
    }
    h8 b1, b2;
    __builtin_memcpy(&b1, v1, sizeof(v1));
    __builtin_memcpy(&b2, v2, sizeof(v2));
#pragma unroll
    for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.a[t], b1, acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.a[t], b2, acc[t], 0, 0, 0);
    if constexpr (Ones) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
            accw[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.a[t], o.b3[0], accw[t], 0, 0, 0);
    }
}

// One sweep into acc (zeroed here): `cur` with its group 0 already in s0; on return s0
// holds group 0 of the sweep `nxt`, whose loads stay in flight through whatever the caller
// does next. Two explicit operand sets with straight-line pairs of groups: each MFMA
// block waits only for its own set while the other set's loads are in flight (the
// scheduling barriers keep each load block ahead of the MFMA block before it).
template <int NT, bool Ones>
__device__ __forceinline__ void run_sweep(mfma_v4 (&acc)[NT], mfma_v4 (&accw)[NT],
                                          const HSweep<NT>& cur, const HSweep<NT>& nxt,
                                          int32_t ng, HSet<NT, Ones>& s0,
                                          HSet<NT, Ones>& s1, int32_t d = 0) {
    const int paddr = a_permute_addr();
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        acc[t] = mfma_v4{0.f, 0.f, 0.f, 0.f};
        if constexpr (Ones) accw[t] = mfma_v4{0.f, 0.f, 0.f, 0.f};
    }
    int32_t g = 0;
    for (; g + 1 < ng; g += 2) {
        h_load(cur, g + 1, s1, d);
        __builtin_amdgcn_sched_barrier(0);
        a_permute(s0, paddr);
        h_mma(acc, accw, s0);
        __builtin_amdgcn_sched_barrier(0);
        // group g + 2 of this sweep, or group 0 of the next one (addresses selected, so
        // the loads are unconditional)
        const bool more = g + 2 < ng;
        HSweep<NT> w;
        w.X = more ? cur.X : nxt.X;
        w.C = more ? cur.C : nxt.C;
#pragma unroll
        for (int t = 0; t < NT; ++t) w.xo[t] = more ? cur.xo[t] : nxt.xo[t];
        w.co = more ? cur.co : nxt.co;
        h_load(w, more ? g + 2 : 0, s0, d);
        __builtin_amdgcn_sched_barrier(0);
        a_permute(s1, paddr);
        h_mma(acc, accw, s1);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (g < ng) {   // an odd group count: the last group is in s0
        a_permute(s0, paddr);
        h_mma(acc, accw, s0);
        __builtin_amdgcn_sched_barrier(0);
        h_load(nxt, 0, s0, d);
    }
}

// ---------------------------------------------------------------------------------------
// Diagonal-sum sweeps over K-group pairs: per tile, the sample operands of two K groups are
// loaded together and run through their MFMAs as one step. Two load layouts (Direct):
//  * Direct (common weight, strides below kQuadQ; per-sample weights, strides below
//    kQuadQWeighted): each lane loads its operand in the MFMA layout itself (row l % 16 of
//    the tile, K slice l / 16: 16 bytes), so one load touches 16 rows of 64 bytes, the 4
//    lanes of a row 16 apart.
//  * Quad (common weight from stride kQuadQ on, per-sample weights from kQuadQWeighted on):
//    a load covers 8 rows x 128 bytes (two K groups): lane l reads row 8 h + (l / 4) % 8 of
//    the tile (h = the tile's load 0 or 1), K group 2 p + l / 32, slice l % 4 (16 bytes), so
//    most 128-byte lines are fetched once. v_permlane32_swap then gathers the first K group of all 16
//    rows into the load-0 register and the second into the load-1 register (lanes
//    32 h + 4 r8 + s: row 8 h + r8, slice s), and ds_bpermute (source lane 4 r + s) moves
//    them to the operand layout, 4 KiB per K group and wavefront through the CU's LDS
//    crossbar.
// kQuadQ is the default of the tunable threshold (gtls_set_tuning).
// Rows are 4 q bytes apart. Measured on the batched workload with a common weight: Direct
// everywhere 6.4 % faster than Quad everywhere (with Quad the LDS was active 63 % of the
// kernel's cycles); Quad from q = 96 on (where the sweeps are long and the operand moves
// few per candidate) another 6.6 % off the kernel's time on batch 0 (32: 2.1 %, 64: 4.7 %,
// 128: 6.4 %, 192: 5.5 %, 256: 4.5 %). With per-sample weights, Direct everywhere is 7.7 %
// slower than Quad everywhere, but Direct below stride 24 is 4.7 % faster on 128 weighted
// curves (bound 16: 5.3 %, 32: 4.8 %, 40: 3.3 %, 64: 2.3 %).
constexpr int kQuadQ = 96;
constexpr int kQuadQWeighted = 24;
// ---------------------------------------------------------------------------------------
struct PSweep {
    const char* X;       // sample pair plane (wave-uniform)
    const char* C;       // template pair plane (wave-uniform)
    int32_t rb;          // lane's row in tile 0 (load 0): 64 st + l % 16, Quad 64 st + (l / 4) % 8
    int32_t co;          // lane's B operand byte offset (K group 0)
    lds_char* LB;        // lane's staged B operand of group 0 (BL)
};

// Super-tile st's pair sweep over X and C: A operand X[row q + b] (samples b >= q meet zero
// template entries and rows past the data read block 0, so the sample plane needs only
// finite values: the gather zero-fills its tail); B operand C[c q_pad + b] for columns
// c <= F, the ones chunks F + 1 and F + 2 in columns 14 and 15 when `ones` (F <= 13), and
// the zero block (zrel bytes from C) in the other columns.
template <bool Direct>
__device__ __forceinline__ PSweep p_sweep(const char* __restrict__ X, const char* __restrict__ C,
                                          int32_t zrel, int32_t q, int32_t F, int64_t st,
                                          bool ones, lds_char* LB) {
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    PSweep w;
    w.X = X;
    w.C = C;
    w.rb = (int32_t)(64 * st) + (Direct ? (lane & 15) : ((lane >> 2) & 7));
    const int32_t ch = col <= F ? col : (ones && col >= kOnesCol ? F + 1 + (col - kOnesCol) : -1);
    w.co = (ch >= 0) ? 4 * (ch * chunk_pad(q) + 4 * kk) : zrel + 16 * kk;
    w.LB = LB;
    return w;
}

// One K-group pair's operands: per tile the two loads, and the B operands of both groups.
struct PSet {
    h8 a[4][2];
    h8 b[2];
};

// Loads pair p (< ng_pairs) of a sweep (rows q samples apart, rows past `rows` read
// block 0); the B operand of a second group past ng - 1 (odd ng) repeats group ng - 1 and
// is not used.
template <bool BL, bool Direct>
__device__ __forceinline__ void p_load(const PSweep& w, int32_t p, int32_t ng, int32_t q,
                                       int32_t rows, PSet& o) {
    const int lane = lane_id();
    if constexpr (Direct) {
        // K slice l / 16 of groups 2 p (h = 0) and 2 p + 1 (h = 1)
        const uint32_t off = 128u * (uint32_t)p + 16u * (uint32_t)(lane >> 4);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int32_t row = w.rb + 16 * t;
            const uint32_t ro = row < rows ? 4u * (uint32_t)row * (uint32_t)q : 0u;
#pragma unroll
            for (int h = 0; h < 2; ++h) o.a[t][h] = ld_h8(w.X + (ro + off + 64u * (uint32_t)h));
        }
    } else {
        const uint32_t off =
            128u * (uint32_t)p + 64u * (uint32_t)(lane >> 5) + 16u * (uint32_t)(lane & 3);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int32_t row = w.rb + 16 * t + 8 * h;
                const uint32_t ro = row < rows ? 4u * (uint32_t)row * (uint32_t)q : 0u;
                o.a[t][h] = ld_h8(w.X + (ro + off));
            }
    }
    const int32_t g0 = 2 * p, g1 = min(2 * p + 1, ng - 1);
    if constexpr (BL) {
        o.b[0] = *(const __attribute__((address_space(3))) h8*)(w.LB + 1024 * g0);
        o.b[1] = *(const __attribute__((address_space(3))) h8*)(w.LB + 1024 * g1);
    } else {
        o.b[0] = ld_h8(w.C + (int64_t)(w.co + 64 * g0));
        o.b[1] = ld_h8(w.C + (int64_t)(w.co + 64 * g1));
    }
}

// The MFMAs of both K groups (the second only when `two`); Quad loads first move the
// operands to the MFMA layout (permlane32 swap, then ds_bpermute from lane paddr / 4).
template <bool Direct>
__device__ __forceinline__ void p_mma(mfma_v4 (&acc)[4], PSet& o, int paddr, bool two) {
    if constexpr (!Direct) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint32_t v0[4], v1[4];
            __builtin_memcpy(v0, &o.a[t][0], sizeof(v0));
            __builtin_memcpy(v1, &o.a[t][1], sizeof(v1));
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const auto sw = __builtin_amdgcn_permlane32_swap(v0[e], v1[e], false, false);
                v0[e] = (uint32_t)__builtin_amdgcn_ds_bpermute(paddr, (int)sw[0]);
                v1[e] = (uint32_t)__builtin_amdgcn_ds_bpermute(paddr, (int)sw[1]);
            }
            __builtin_memcpy(&o.a[t][0], v0, sizeof(v0));
            __builtin_memcpy(&o.a[t][1], v1, sizeof(v1));
        }
    }
    HSet<4> g;
    g.b = o.b[0];
#pragma unroll
    for (int t = 0; t < 4; ++t) g.a[t] = o.a[t][0];
    mfma_v4 unused[4];
    h_mma(acc, unused, g);
    if (two) {
        g.b = o.b[1];
#pragma unroll
        for (int t = 0; t < 4; ++t) g.a[t] = o.a[t][1];
        h_mma(acc, unused, g);
    }
}

// One pair sweep into acc (zeroed here), pipelined as run_sweep over pairs: `cur` with its
// pair 0 already in s0; on return s0 holds pair 0 of `nxt`.
template <bool BL, bool Direct>
__device__ __forceinline__ void run_psweep(mfma_v4 (&acc)[4], const PSweep& cur,
                                           const PSweep& nxt, int32_t ng, int32_t q,
                                           int32_t rows, PSet& s0, PSet& s1) {
    const int paddr = a_permute_addr();
    const int32_t np = (ng + 1) >> 1;
    const bool odd = (ng & 1) != 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma_v4{0.f, 0.f, 0.f, 0.f};
    int32_t p = 0;
    for (; p + 1 < np; p += 2) {
        p_load<BL, Direct>(cur, p + 1, ng, q, rows, s1);
        __builtin_amdgcn_sched_barrier(0);
        p_mma<Direct>(acc, s0, paddr, true);
        __builtin_amdgcn_sched_barrier(0);
        // pair p + 2 of this sweep, or pair 0 of the next one
        const bool more = p + 2 < np;
        PSweep w;
        w.X = more ? cur.X : nxt.X;
        w.C = more ? cur.C : nxt.C;
        w.rb = more ? cur.rb : nxt.rb;
        w.co = more ? cur.co : nxt.co;
        w.LB = more ? cur.LB : nxt.LB;
        p_load<BL, Direct>(w, more ? p + 2 : 0, ng, q, rows, s0);
        __builtin_amdgcn_sched_barrier(0);
        p_mma<Direct>(acc, s1, paddr, !(odd && p + 2 == np));
        __builtin_amdgcn_sched_barrier(0);
    }
    if (p < np) {   // an odd pair count: the last pair is in s0
        p_mma<Direct>(acc, s0, paddr, !odd);
        __builtin_amdgcn_sched_barrier(0);
        p_load<BL, Direct>(nxt, 0, ng, q, rows, s0);
    }
}

// The pair sweep `cur` of a super-tile, written to the ring (see run_psweep).
template <bool BL, bool Direct>
__device__ __forceinline__ void super_tile_prows(const PSweep& cur, const PSweep& nxt, int32_t ng,
                                                 int32_t q, int32_t rows, PSet& s0, int ring0,
                                                 float* ring) {
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    mfma_v4 acc[4];
    PSet s1;
    run_psweep<BL, Direct>(acc, cur, nxt, ng, q, rows, s0, s1);
    // rows 16 t + 4 kk .. + 3 of column col: one 16-byte store (ring0 is a multiple of 16
    // and kRingRows of 4, so the four rows never wrap apart), again at the mirror rows
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int r = ring0 + 16 * t + 4 * kk;
        r = r >= kRingRows ? r - kRingRows : r;
        *(mfma_v4*)(ring + col * kRingPitch + r) = acc[t];
        if (r < kRingMirror) *(mfma_v4*)(ring + col * kRingPitch + kRingRows + r) = acc[t];
    }
}

// LDS of each wavefront of a workgroup in the diagonal-sum path (a quarter of 80 KiB, two
// workgroups per CU): the ring of A sums, then either the staged template operands
// (common weight) or the ring of B sums and the staged operands of both planes. Staged
// operands: for each K group g of the width, the B operand of every lane l at 1024 g +
// 16 l (S plane; the S^2 plane follows after ng groups).
constexpr int kWaveLdsBytes = 20480;
constexpr int kRingBytes = 4 * kRingFloats;
constexpr int kStageGroupsConstW = (kWaveLdsBytes - kRingBytes) / 1024;        // 13
constexpr int kStageGroupsWeighted = (kWaveLdsBytes - 2 * kRingBytes) / 2048;  // 3
__shared__ __attribute__((aligned(16))) float g_wave_lds[4 * kWaveLdsBytes / 4];

// Sample pair planes of one (period, width) item of the matrix-core paths: the A-sum
// samples X (u with a common weight, v = u w otherwise) and the weights W (weighted
// path), and 1 / scale of each.
struct HPlanes {
    const char* X;
    const char* W;
    float inv_x, inv_w;
};

// Every candidate of one (period, width) item on the matrix cores, returning the lane's
// first minimum of the statistic. C, C2: chunked template pair planes of S and S^2
// (their zero blocks zrel bytes from C and C2). Reads of X stay below N + d + q + 16
// (zero-filled past the gathered samples). The window means: with TileMean (common
// weight, F <= 13), from the ones columns of the same tiles, the window sum of
// u = sum_{c<F} G(k + c, 14) + G(k + F, 15) in float32; otherwise from D, the float64
// prefix sums of u. With BL the template operands of every K group are staged once in the
// wavefront's LDS (ng <= kStageGroupsConstW, or kStageGroupsWeighted for both planes)
// instead of being read from memory by every super-tile.
// FC: the chunk count F = d / q when known at compile time (8: every width with q >= 8 at
// the default skip of 8), so that the diagonal reads take immediate offsets and the columns
// past F are not read at all; 0 for any F. Direct: the sample operands' load layout (see
// p_load; Quad when false).
template <bool ConstW, bool TileMean, bool BL, int FC = 0, bool Direct = ConstW>
__device__ __forceinline__ ArgMin mfma_width(const HPlanes& P, const char* __restrict__ C,
                                             const char* __restrict__ C2, int32_t zrel,
                                             const double* __restrict__ D,
                                             WidthArgs a) {
    static_assert(ConstW || !TileMean, "tile means need the samples u");
    const int lane = lane_id();
    const int wave = uniform((int32_t)wave_id());
    float* ringA = g_wave_lds + wave * (kWaveLdsBytes / 4);
    float* ringB = ringA + kRingFloats;
    const int32_t d = a.d, q = a.q;
    const int64_t K = a.K;
    const int32_t F = FC ? FC : d / q;
    const int64_t rows = K + F;
    const int64_t nsuper = (rows + 63) / 64;
    LaneMin L = lane_min_init();
    const WidthConstsF fc = width_consts_f(a);
    const float inv_xd = P.inv_x * fc.inv_d;
    const int32_t ng = chunk_pad(q) / kKGroup;
    // the staged operands: after the ring(s); lane's slot of group 0 of each plane
    lds_char* stage = (lds_char*)(ringA + (ConstW ? 1 : 2) * kRingFloats) + 16 * lane;
    auto sweep = [&](const char* Xp, const char* Cp, int64_t st, bool ones, int plane) {
        return p_sweep<Direct>(Xp, Cp, zrel, q, F, st, ones, stage + 1024 * ng * plane);
    };
    if constexpr (BL) {
        // every lane's B operand of each K group (and plane), 4 groups of loads in flight
        for (int pl = 0; pl < (ConstW ? 1 : 2); ++pl) {
            const PSweep w = sweep(P.X, pl ? C2 : C, 0, TileMean, pl);
            const char* src = w.C + w.co;
            for (int32_t g0 = 0; g0 < ng; g0 += 4) {
                h8 v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = ld_h8(src + 64 * min(g0 + e, ng - 1));
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (g0 + e < ng)
                        *(__attribute__((address_space(3))) h8*)(w.LB + 1024 * (g0 + e)) = v[e];
            }
        }
    }
    int ring0 = 0;   // ring row of G row 64 T
    // sweeps in order (A plane, then B plane, of each super-tile); the first group of each
    // sweep is loaded during the one before (past the last super-tile: row block 0 again)
    PSet ops;
    const int32_t rows32 = (int32_t)rows;
    p_load<BL, Direct>(sweep(P.X, C, 0, TileMean, 0), 0, ng, q, rows32, ops);
    for (int64_t st = 0; st < nsuper; ++st) {
        const int32_t kc = (int32_t)(64 * st) - F + lane;   // K <= N <= 2^30
        const bool cand = kc >= 0 && kc < (int32_t)K;
        // the window-mean prefix sums of this super-tile's candidates, requested early
        double dlo = 0.0, dhi = 0.0;
        if constexpr (!TileMean) {
            if (cand) {
                const int64_t s = kc * q;
                dhi = D[s + d - 1];
                dlo = s ? D[s - 1] : 0.0;
            }
        }
        const int64_t stn = st + 1 < nsuper ? st + 1 : 0;
        const PSweep wa = sweep(P.X, C, st, TileMean, 0);
        if constexpr (ConstW) {
            super_tile_prows<BL, Direct>(wa, sweep(P.X, C, stn, TileMean, 0), ng, q, rows32, ops,
                                         ring0, ringA);
        } else {
            const PSweep wb = sweep(P.W, C2, st, false, 1);
            super_tile_prows<BL, Direct>(wa, wb, ng, q, rows32, ops, ring0, ringA);
            super_tile_prows<BL, Direct>(wb, sweep(P.X, C, stn, false, 0), ng, q, rows32, ops,
                                         ring0, ringB);
        }
        wave_sync_lds();
        // diagonal sums: column c at the lane's row rw + min(c, F) (rows rw .. rw + F, this
        // super-tile's and the previous one's, mirrored: no wrap), all reads issued before
        // the first add. Columns F + 1 .. 15 of the A and B sums met the zero block, so
        // their entries in row rw + F are exact zeros and are added unmasked (the ones
        // columns 14 and 15 of TileMean are left out): the same terms in the same order.
        // Rows past rw + F can still hold the previous item's values; the block sums read
        // there (c >= F) are masked.
        constexpr int NA = FC ? FC + 1 : (TileMean ? kOnesCol : 16);   // columns read
        constexpr int NU = kOnesCol - 2;   // TileMean: F <= 12 block sums
        const int rw = ring_wrap(ring0 - F + lane);
        const float* ra = ringA + rw;
        const float* rb = ringB + rw;
        float va[NA], vb[ConstW ? 1 : 16], vu[TileMean ? NU + 1 : 1];
#pragma unroll
        for (int c = 0; c < NA; ++c) va[c] = ra[c * kRingPitch + min(c, F)];
        if constexpr (!ConstW) {
#pragma unroll
            for (int c = 0; c < 16; ++c) vb[c] = rb[c * kRingPitch + min(c, F)];
        }
        if constexpr (TileMean) {
            // block sums (column 14) of blocks k .. k + F - 1, the rem sum (15) of block k + F
#pragma unroll
            for (int c = 0; c < NU; ++c) vu[c] = ra[kOnesCol * kRingPitch + c];
            vu[NU] = ra[(kOnesCol + 1) * kRingPitch + F];
        }
        float sa = 0.f, sb = 0.f, su = 0.f;
#pragma unroll
        for (int c = 0; c < NA; ++c) sa += va[c];
        if constexpr (!ConstW) {
#pragma unroll
            for (int c = 0; c < 16; ++c) sb += vb[c];
        }
        if constexpr (TileMean) {
#pragma unroll
            for (int c = 0; c < NU; ++c) su += c < F ? vu[c] : 0.f;
            su += vu[NU];
        }
        sa *= P.inv_x;
        if constexpr (!ConstW) sb *= P.inv_w;
        wave_sync_lds();
        if (cand) {
            const int32_t s = kc * q;
            if constexpr (TileMean) {
                const float m = su * inv_xd;
                tile_candidate<ConstW>(L, a, fc, m > fc.depth_min, m * fc.os2, sa, sb, s);
            } else {
                const double m = (dhi - dlo) * a.inv_d;
                tile_candidate<ConstW>(L, a, fc, m > a.depth_min, (float)(m * a.os2), sa, sb, s);
            }
        }
        ring0 = ring0 + 64 >= kRingRows ? ring0 + 64 - kRingRows : ring0 + 64;
    }
    return L.best;
}

// ---------------------------------------------------------------------------------------
// Short strides (q < kToepQ) on the matrix cores as Toeplitz tiles: with rows i of a tile
// starting at sample q (256 T + 16 i) and column j holding the template shifted by q j,
//     C(i, j) = sum_k X[q (256 T + 16 i) + k] S[k - q j] = A(candidate 256 T + 16 i + j),
// k running over the d + 15 q taps where a column can be non-zero (in K groups of 16). No
// diagonal sums: the result register v of lane l is candidate 256 T + 16 (4 (l / 16) + v)
// + l % 16. NT tiles (256 NT candidates, a step) share each template operand. The
// template rows carry kTmplPad (>= 15 q + 16) zeros on both sides, so taps past a shifted
// template read zeros, and the sample plane is zero past the data (reads stay below
// N + d + 15 q + 16). T, T2: template pair planes of S and S^2. With a common weight the window sums of u come from
// the same tiles with a row of d ones (formed in registers, HSet) in float32; with
// per-sample weights from D, the float64 prefix sums of u.
// ---------------------------------------------------------------------------------------
constexpr int kToepQ = 24;        // largest stride bound the template padding allows
constexpr int kToepQDefault = 8;  // measured best on the batched workload (24: +2.5 %)
constexpr int kToepTiles = 4;
constexpr int kToepStep = 256 * kToepTiles;   // candidates per step

template <bool ConstW>
__device__ __forceinline__ ArgMin toep_width(const HPlanes& P, const char* __restrict__ T,
                                             const char* __restrict__ T2,
                                             const double* __restrict__ D, WidthArgs a) {
    constexpr int NT = kToepTiles;
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    const int32_t d = a.d, q = a.q;
    const int64_t K = a.K;
    const int32_t ng = k_groups(d + 15 * q);         // d + 15 q taps
    const int64_t nsuper = (K + kToepStep - 1) / kToepStep;
    LaneMin L = lane_min_init();
    const WidthConstsF fc = width_consts_f(a);
    const float inv_xd = P.inv_x * fc.inv_d;
    // sweep of step st over one plane: row `col` of tile t covers candidates
    // kToepStep st + 256 t + 16 col .. + 15 (rows wholly past the last candidate read
    // sample 0, results unused); B operand: template pairs at taps 4 kk .. + 3 shifted by
    // q col
    const int qr = lane >> 2, qk = lane & 3;   // load lanes: row qr, K slice qk (a_permute)
    auto sweep = [&](const char* Xp, const char* Tp, int64_t st) {
        HSweep<NT> w;
        w.X = Xp;
        w.C = Tp;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int64_t m0 = kToepStep * st + 256 * t + 16 * qr;
            w.xo[t] = (uint32_t)(4 * ((m0 < K ? m0 * q : 0) + 4 * qk));
        }
        w.co = 4 * (4 * kk - q * col);
        return w;
    };
    mfma_v4 acc[NT], acc2[NT];
    HSet<NT, ConstW> s0, s1;
    h_load(sweep(P.X, T, 0), 0, s0, d);
    for (int64_t st = 0; st < nsuper; ++st) {
        const int64_t stn = st + 1 < nsuper ? st + 1 : 0;
        // the window-mean prefix sums D[s + d - 1] and D[s - 1] (0 at s = 0) of the lane's
        // 4 NT candidates (kToepStep st + 256 t + 16 (4 kk + v) + col), requested before
        // the sweeps so that they arrive while the MFMAs run
        double dlo[ConstW ? 1 : 4 * NT], dhi[ConstW ? 1 : 4 * NT];
        if constexpr (!ConstW) {
#pragma unroll
            for (int e = 0; e < 4 * NT; ++e) {
                const int64_t m = kToepStep * st + 256 * (e >> 2) + 16 * (4 * kk + (e & 3)) + col;
                const int64_t s = (m < K ? m : 0) * q;
                dhi[e] = D[s + d - 1];
                dlo[e] = D[s > 0 ? s - 1 : 0];
            }
        }
        // A sums (and the sums of u with a common weight, the B sums with the weights); the
        // first group of the next sweep is loaded during this one
        if constexpr (ConstW) {
            run_sweep<NT, true>(acc, acc2, sweep(P.X, T, st), sweep(P.X, T, stn), ng, s0, s1,
                                d);
        } else {
            run_sweep<NT, false>(acc, acc2, sweep(P.X, T, st), sweep(P.W, T2, st), ng, s0, s1);
            mfma_v4 unused[NT];
            run_sweep<NT, false>(acc2, unused, sweep(P.W, T2, st), sweep(P.X, T, stn), ng, s0,
                                 s1);
        }
        // the statistic of the lane's 4 NT candidates (indices and positions fit int32:
        // K <= N <= 2^30)
        const int32_t mbase = (int32_t)(kToepStep * st) + 64 * kk + col;
        const int32_t K32 = (int32_t)K;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int e = 4 * t + v;
                const int32_t m = mbase + 256 * t + 16 * v;
                if (m < K32) {
                    const int32_t s = m * q;
                    const float sa = acc[t][v] * P.inv_x;
                    if constexpr (ConstW) {
                        const float mm = acc2[t][v] * inv_xd;
                        tile_candidate<true>(L, a, fc, mm > fc.depth_min, mm * fc.os2, sa, 0.f, s);
                    } else {
                        const double mm = (dhi[e] - (s > 0 ? dlo[e] : 0.0)) * a.inv_d;
                        tile_candidate<false>(L, a, fc, mm > a.depth_min, (float)(mm * a.os2), sa,
                                              acc2[t][v] * P.inv_w, s);
                    }
                }
            }
        }
    }
    return L.best;
}

}  // namespace gtls
