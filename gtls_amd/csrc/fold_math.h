// The fold quotient x = t / P, correctly rounded, from a reciprocal computed once per
// period. Plain C++ (host and device), so that host tests can compile it.
//
// With rP = RN(1 / P) (one correctly rounded division per period):
//   q1 = RN(t rP)                       within about 2 ulps of t / P,
//   q2 = RN(q1 + RN(t - q1 P) rP)       within 1/2 ulp plus a few 2^-53 ulps (faithful),
//   q3 = RN(q2 + (t - q2 P) rP)         the correctly rounded t / P,
// the last step by Markstein's theorem (IBM J. Res. Dev. 34(1), 1990; Muller et al.,
// Handbook of Floating-Point Arithmetic, "Markstein's theorem" on the correction step):
// if y is within 1/2 ulp of 1/b and q within 1 ulp of a/b, then r = a - b q is exact
// (one fused multiply-add) and RN(q + r y) = RN(a / b). Five float64 operations replace
// the division sequence (scale, reciprocal, two Newton steps, fix-ups) per sample.
//
// Where it applies: P in [2^-1000, 2^1000] (fold_reciprocal returns NaN otherwise) and
// the remainder t - q P must not underflow, which holds unless 0 < |t| < 2^-960
// (fold_time_small). Within that, the quotient equals t / P bit for bit whenever t / P is
// finite; the phase x - trunc(x) of the fold also agrees for t = -0 (quotient +0 against
// -0, phase +0 both), infinite or NaN times and overflowing quotients (NaN phases both).
// Callers that meet a small time or a NaN reciprocal divide instead.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define GTLS_FOLD_FN __host__ __device__ __forceinline__
#else
#define GTLS_FOLD_FN static inline
#endif

namespace gtls {

// RN(1 / P), or NaN where the reciprocal form does not apply
GTLS_FOLD_FN double fold_reciprocal(double P) {
    const double a = fabs(P);
    return (a >= 0x1p-1000 && a <= 0x1p1000) ? 1.0 / P : __builtin_nan("");
}

// 0 < |t| < 2^-960: the reciprocal form could misround t / P
GTLS_FOLD_FN bool fold_time_small(double t) {
    uint64_t b;
    memcpy(&b, &t, sizeof b);
    const uint64_t mag = b & 0x7fffffffffffffffULL;
    return mag != 0 && mag < 0x03f0000000000000ULL;   // 0x03f0... is 2^-960
}

GTLS_FOLD_FN double fold_quotient(double t, double P, double rP) {
    double q = t * rP;
    double r = fma(-q, P, t);
    q = fma(r, rP, q);
    r = fma(-q, P, t);
    return fma(r, rP, q);
}

}  // namespace gtls
