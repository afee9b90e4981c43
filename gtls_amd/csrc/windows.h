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
// current one, and the first group of the next sweep (the B plane, or the next
// super-tile) before the last MFMAs of this one, so its latency passes during the
// diagonal sums.
// ---------------------------------------------------------------------------------------
typedef float mfma_v4 __attribute__((ext_vector_type(4)));
constexpr int kRingRows = 80;     // 64 rows of a super-tile + up to 15 look-ahead rows
constexpr int kRingPitch = 84;    // column-major: column c at c * kRingPitch (16-byte rows of 4)
constexpr int kRingFloats = 16 * kRingPitch;


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

__device__ __forceinline__ int ring_wrap(int r) {
    r = r < 0 ? r + kRingRows : r;
    return r >= kRingRows ? r - kRingRows : r;
}


// K-group size of the matrix-core path: 4 M samples per group, M MFMA steps per tile. Wide
// strides take long groups so that the MFMAs of one group cover the latency of the next
// group's loads -- unless 32-sample groups would pad q further than 16-sample ones;
// short strides take short groups so that little of a group is padding.
__host__ __device__ inline int32_t mfma_m(int32_t q) {
    if (q >= 48) return (q + 31) / 32 * 32 == (q + 15) / 16 * 16 ? 8 : 4;
    return q >= 16 ? 4 : 2;
}

// Chunked template rows of the matrix-core path: for width w with stride q, F = d / q and
// q_pad = q rounded up to whole K groups (4 mfma_m(q) samples), chunk c <= F sits at
// chunk_offset(w) + c q_pad, holds S[c q + b] for b < its length (q, or d - F q for
// c = F) and zeros up to q_pad. Columns c > F read a block of zeros instead.
// (F + 1) q_pad <= 2 d + 1024 for F <= 15, so the offsets below leave room for every row.
__host__ __device__ inline int64_t chunk_offset(int64_t tmpl_off, int64_t w) {
    return 2 * tmpl_off + 1024 * w;
}
__host__ __device__ inline int32_t chunk_pad(int32_t q) {
    const int32_t g = 4 * mfma_m(q);
    return (q + g - 1) / g * g;
}

// Operand addresses of one sweep (the 64 G rows of super-tile st for one sample plane X
// and chunked template C): the lane's A pointers (4 tiles) and B pointer, and the group
// count. A operand X[row q + b]: samples b >= q meet zero template entries and rows past
// the data read block 0, so the sample plane needs only finite values (the gather
// zero-fills its tail); B operand C[c q_pad + b], columns c > F read the zero block Z.
template <int M>
struct GSweep {
    const float* X;      // sample plane (wave-uniform)
    const float* C;      // template rows (wave-uniform)
    uint32_t xo[4];      // lane's A operand byte offsets from X, one per tile
    int32_t co;          // lane's B operand byte offset from C
};

template <int M>
__device__ __forceinline__ GSweep<M> g_sweep(const float* __restrict__ X,
                                             const float* __restrict__ C,
                                             const float* __restrict__ Z, int32_t q, int32_t F,
                                             int64_t rows, int64_t st) {
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    GSweep<M> w;
    w.X = X;
    w.C = C;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int64_t row = 64 * st + 16 * t + col;
        w.xo[t] = (uint32_t)(4 * ((row < rows ? row * q : 0) + M * kk));
    }
    w.co = (int32_t)(4 * ((col <= F) ? (int64_t)col * chunk_pad(q) + M * kk
                                     : (int64_t)(Z - C) + M * kk));
    return w;
}

// One K group's operands of a sweep (M MFMA steps for each of the 4 tiles).
template <int M>
struct GSet {
    FV<M> a[4];
    FV<M> b;
};

template <int M>
__device__ __forceinline__ void g_load(const GSweep<M>& w, int32_t g, GSet<M>& o) {
    const uint32_t off = 16 * M * g;
    const char* x = (const char*)w.X;
    const char* c = (const char*)w.C;
#pragma unroll
    for (int t = 0; t < 4; ++t) o.a[t] = ldm<M>((const float*)(x + (w.xo[t] + off)));
    o.b = ldm<M>((const float*)(c + (int64_t)(w.co + (int32_t)off)));
}

// K groups of a sweep over `len` taps (4 M samples per group; taps past `len` meet zero
// template entries).
__device__ __forceinline__ int32_t k_groups(int32_t len, int M) {
    return (len + 4 * M - 1) / (4 * M);
}

// One sweep into acc (zeroed here): `cur` with its group 0 already in s0; on return s0
// holds group 0 of the sweep `nxt`, whose loads stay in flight through whatever the caller
// does next. Two explicit operand sets with straight-line pairs of groups: each MFMA
// block waits only for its own set while the other set's loads are in flight (the
// scheduling barriers keep each load block ahead of the MFMA block before it). Within a
// group, step i takes sample 4 M g + M kk + i in K slot kk, so each lane's operands are M
// consecutive floats.
template <int M>
__device__ __forceinline__ void run_sweep(mfma_v4 (&acc)[4], const GSweep<M>& cur,
                                          const GSweep<M>& nxt, int32_t ng, GSet<M>& s0,
                                          GSet<M>& s1) {
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma_v4{0.f, 0.f, 0.f, 0.f};
    auto mma = [&](const GSet<M>& o) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[t].v[i], o.b.v[i], acc[t],
                                                             0, 0, 0);
        }
    };
    int32_t g = 0;
    for (; g + 1 < ng; g += 2) {
        g_load(cur, g + 1, s1);
        __builtin_amdgcn_sched_barrier(0);
        mma(s0);
        __builtin_amdgcn_sched_barrier(0);
        // group g + 2 of this sweep, or group 0 of the next one (addresses selected, so
        // the loads are unconditional)
        const bool more = g + 2 < ng;
        GSweep<M> w;
        w.X = more ? cur.X : nxt.X;
        w.C = more ? cur.C : nxt.C;
#pragma unroll
        for (int t = 0; t < 4; ++t) w.xo[t] = more ? cur.xo[t] : nxt.xo[t];
        w.co = more ? cur.co : nxt.co;
        g_load(w, more ? g + 2 : 0, s0);
        __builtin_amdgcn_sched_barrier(0);
        mma(s1);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (g < ng) {   // an odd group count: the last group is in s0
        mma(s0);
        __builtin_amdgcn_sched_barrier(0);
        g_load(nxt, 0, s0);
    }
}

// The sweep `cur` of a super-tile, written to the ring (see run_sweep).
template <int M>
__device__ __forceinline__ void super_tile_rows(const GSweep<M>& cur, const GSweep<M>& nxt,
                                                int32_t ng, GSet<M>& s0, int ring0,
                                                float* ring) {
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    mfma_v4 acc[4];
    GSet<M> s1;
    run_sweep<M>(acc, cur, nxt, ng, s0, s1);
    // rows 16 t + 4 kk .. + 3 of column col: one 16-byte store (ring0 is a multiple of 16
    // and kRingRows of 4, so the four rows never wrap apart)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int r = ring0 + 16 * t + 4 * kk;
        r = r >= kRingRows ? r - kRingRows : r;
        *(mfma_v4*)(ring + col * kRingPitch + r) = acc[t];
    }
}

// The window rings of the four wavefronts of a workgroup (A sums, then B sums).
__shared__ __attribute__((aligned(16))) float g_rings[2 * 4 * kRingFloats];

// Every candidate of one (period, width) item on the matrix cores, returning the lane's
// first minimum of the statistic. X, C: the sample plane and chunked template of the A
// sums (u with S when ConstW, v = u w with S otherwise); X2, C2: those of the B sums
// (w with S^2; unused when ConstW); Z: zeros (at least q_pad floats). Reads of X stay
// below N + d + q + 64 (zero-filled past the gathered samples). D: prefix sums of u
// (float64) for the window means.
template <bool ConstW, int M>
__device__ __forceinline__ ArgMin mfma_width(const float* __restrict__ X,
                                          const float* __restrict__ X2,
                                          const float* __restrict__ C,
                                          const float* __restrict__ C2,
                                          const float* __restrict__ Z,
                                          const double* __restrict__ D, WidthArgs a) {
    const int lane = lane_id();
    const int wave = uniform((int32_t)wave_id());
    float* ringA = g_rings + wave * kRingFloats;
    float* ringB = g_rings + (4 + wave) * kRingFloats;
    const int32_t d = a.d, q = a.q;
    const int64_t K = a.K;
    const int32_t F = d / q;
    const int64_t rows = K + F;
    const int64_t nsuper = (rows + 63) / 64;
    ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};
    const int32_t ng = k_groups(q, M);
    int ring0 = 0;   // ring row of G row 64 T
    // sweeps in order (A plane, then B plane, of each super-tile); the first group of each
    // sweep is loaded during the one before (past the last super-tile: row block 0 again)
    GSet<M> ops;
    g_load(g_sweep<M>(X, C, Z, q, F, rows, 0), 0, ops);
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
        const int64_t stn = st + 1 < nsuper ? st + 1 : 0;
        const GSweep<M> wa = g_sweep<M>(X, C, Z, q, F, rows, st);
        if constexpr (ConstW) {
            super_tile_rows<M>(wa, g_sweep<M>(X, C, Z, q, F, rows, stn), ng, ops, ring0, ringA);
        } else {
            const GSweep<M> wb = g_sweep<M>(X2, C2, Z, q, F, rows, st);
            super_tile_rows<M>(wa, wb, ng, ops, ring0, ringA);
            super_tile_rows<M>(wb, g_sweep<M>(X, C, Z, q, F, rows, stn), ng, ops, ring0, ringB);
        }
        wave_sync_lds();
        // diagonal sums: all 16 reads issue before the first add, columns past F masked
        float va[16], vb[16];
        const int r0 = ring0 - F + lane;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int r = ring_wrap(r0 + min(c, F));
            va[c] = ringA[c * kRingPitch + r];
            if constexpr (!ConstW) vb[c] = ringB[c * kRingPitch + r];
        }
        float sa = 0.f, sb = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            sa += c <= F ? va[c] : 0.f;
            if constexpr (!ConstW) sb += c <= F ? vb[c] : 0.f;
        }
        wave_sync_lds();
        if (cand) {
            const int32_t s = (int32_t)(kc * q);
            const double m = (dhi - dlo) * a.inv_d;
            if (m > a.depth_min) {
                const float rs = (float)(m * a.os2);
                const double delta =
                    ConstW ? a.wc * ((double)(rs * rs) * a.ssk - 2.0 * (double)rs * (double)sa)
                           : (double)(rs * rs * sb - 2.f * rs * sa);
                consider(best, a.chi2f + delta, a.k, s);
            } else {
                consider(best, a.sentinel, a.k, s);
            }
        }
        ring0 = ring0 + 64 >= kRingRows ? ring0 + 64 - kRingRows : ring0 + 64;
    }
    return best;
}

// ---------------------------------------------------------------------------------------
// Short strides (q < kToepQ) on the matrix cores as Toeplitz tiles: with rows i of a tile
// starting at sample q (256 T + 16 i) and column j holding the template shifted by q j,
//     C(i, j) = sum_k X[q (256 T + 16 i) + k] S[k - q j] = A(candidate 256 T + 16 i + j),
// k running over the d + 15 q taps where a column can be non-zero. No diagonal sums: the
// result register v of lane l is candidate 256 T + 16 (4 (l / 16) + v) + l % 16. Four
// tiles (1024 candidates) share each template operand. The template rows carry kTmplPad
// (>= 15 q + 32) zeros on both sides, so taps past a shifted template read zeros, and the
// sample plane is zero past the data (reads stay below N + d + 15 q + 32).
// ---------------------------------------------------------------------------------------
constexpr int kToepQ = 24;

template <bool ConstW>
__device__ __forceinline__ ArgMin toep_width(const float* __restrict__ X,
                                             const float* __restrict__ X2,
                                             const float* __restrict__ T,
                                             const float* __restrict__ T2,
                                             const double* __restrict__ D, WidthArgs a) {
    constexpr int M = 4;
    const int lane = lane_id();
    const int kk = lane >> 4, col = lane & 15;
    const int32_t d = a.d, q = a.q;
    const int64_t K = a.K;
    const int32_t ng = k_groups(d + 15 * q, M);         // d + 15 q taps
    const int64_t nsuper = (K + 1023) / 1024;
    ArgMin best{__builtin_inf(), 0x7fffffff, 0x7fffffff};
    // sweep of step st over one plane: row `col` of tile t covers candidates
    // 1024 st + 256 t + 16 col .. + 15 (rows wholly past the last candidate read sample 0,
    // results unused); B operand: template value at tap 4 M kk + s shifted by q col
    auto sweep = [&](const float* P, const float* Tp, int64_t st) {
        GSweep<M> w;
        w.X = P;
        w.C = Tp;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t m0 = 1024 * st + 256 * t + 16 * col;
            w.xo[t] = (uint32_t)(4 * ((m0 < K ? m0 * q : 0) + M * kk));
        }
        w.co = 4 * (M * kk - q * col);
        return w;
    };
    mfma_v4 acc[4], acc2[4];
    GSet<M> s0, s1;
    g_load(sweep(X, T, 0), 0, s0);
    for (int64_t st = 0; st < nsuper; ++st) {
        const int64_t stn = st + 1 < nsuper ? st + 1 : 0;
        // the window-mean prefix sums D[s + d - 1] and D[s - 1] (0 at s = 0) of the lane's
        // 16 candidates (1024 st + 256 t + 16 (4 kk + v) + col), requested before the
        // sweeps so that they arrive while the MFMAs run
        double dlo[16], dhi[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int64_t m = 1024 * st + 256 * (e >> 2) + 16 * (4 * kk + (e & 3)) + col;
            const int64_t s = (m < K ? m : 0) * q;
            dhi[e] = D[s + d - 1];
            dlo[e] = D[s > 0 ? s - 1 : 0];
        }
        // A sums (and B sums with the weights); the first group of the next sweep is
        // loaded during this one
        if constexpr (ConstW) {
            run_sweep<M>(acc, sweep(X, T, st), sweep(X, T, stn), ng, s0, s1);
        } else {
            run_sweep<M>(acc, sweep(X, T, st), sweep(X2, T2, st), ng, s0, s1);
            run_sweep<M>(acc2, sweep(X2, T2, st), sweep(X, T, stn), ng, s0, s1);
        }
        // the statistic of the lane's 16 candidates
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int e = 4 * t + v;
                const int64_t m = 1024 * st + 256 * t + 16 * (4 * kk + v) + col;
                if (m < K) {
                    const int32_t s = (int32_t)(m * q);
                    const double mm = (dhi[e] - (s > 0 ? dlo[e] : 0.0)) * a.inv_d;
                    if (mm > a.depth_min) {
                        const float rs = (float)(mm * a.os2);
                        const float sa = acc[t][v];
                        const double delta =
                            ConstW ? a.wc * ((double)(rs * rs) * a.ssk - 2.0 * (double)rs * (double)sa)
                                   : (double)(rs * rs * acc2[t][v] - 2.f * rs * sa);
                        consider(best, a.chi2f + delta, a.k, s);
                    } else {
                        consider(best, a.sentinel, a.k, s);
                    }
                }
            }
        }
    }
    return best;
}

// Dispatches on the weight mode and the K-group size of stride q (mfma_m).
template <bool ConstW>
__device__ __forceinline__ ArgMin window_sums_mfma(const float* X, const float* X2,
                                                   const float* C, const float* C2,
                                                   const float* Z, const double* D,
                                                   const WidthArgs& a) {
    const int32_t m = mfma_m(a.q);
    if (m == 8) return mfma_width<ConstW, 8>(X, X2, C, C2, Z, D, a);
    if (m == 4) return mfma_width<ConstW, 4>(X, X2, C, C2, Z, D, a);
    return mfma_width<ConstW, 2>(X, X2, C, C2, Z, D, a);
}

}  // namespace gtls
