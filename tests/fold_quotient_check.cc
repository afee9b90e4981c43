// Host check of gtls::fold_quotient (gtls_amd/csrc/fold_math.h) against IEEE float64
// division, bit for bit, and of the fold phases it gives on special inputs. Usage: fold_quotient_check N_RANDOM. Prints the counts of
// mismatches per case family and of the plain product t * RN(1/P)'s misroundings (which
// must be nonzero: the sample then contains inputs on which a missing correction step
// shows), and exits nonzero on any mismatch.
#include "fold_math.h"

#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>

static bool same(double a, double b) { return memcmp(&a, &b, sizeof a) == 0; }

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 g(20261204);
    std::uniform_real_distribution<double> ut(-3000.0, 12000.0), ue(-3.0, 3.0);
    long bad_random = 0, bad_mid = 0, bad_cadence = 0, bad_special = 0, naive_wrong = 0;
    for (long i = 0; i < n; ++i) {
        const double t = ut(g), P = pow(10.0, ue(g)), rP = gtls::fold_reciprocal(P);
        const double q = t / P;
        if (!same(gtls::fold_quotient(t, P, rP), q)) ++bad_random;
        if (!same(t * rP, q)) ++naive_wrong;
        // t / P within a tiny fraction of an ulp of a rounding midpoint of q
        const double ulp = nextafter(fabs(q), INFINITY) - fabs(q);
        const __float128 mid = (__float128)q + (__float128)(q >= 0 ? ulp : -ulp) / 2;
        const double t2 = (double)(mid * (__float128)P);
        if (!same(gtls::fold_quotient(t2, P, rP), t2 / P)) ++bad_mid;
    }
    // 2-minute cadence times against a fine period grid
    for (long k = 0; k < n / 4; ++k) {
        const double t = 1325.0 + (double)k / 720.0, P = 0.5 + (double)(k % 99991) * 1e-5;
        if (!same(gtls::fold_quotient(t, P, gtls::fold_reciprocal(P)), t / P)) ++bad_cadence;
    }
    // zeros, exact multiples, powers of two, tiny and subnormal times, overflowing ratios,
    // periods at and past the reciprocal form's range: where the form applies (not a small
    // time, a finite reciprocal) the fold's phase x - trunc(x) equals the divided one's
    // (NaN when both are NaN)
    const double ts[] = {0.0, -0.0, 1.0, -1.0, 3.0, 0x1p-960, -0x1p-960, 0x1.8p-961,
                         0x1p-1000, 0x1p-1074, -0x1p-1060, 0x1.fffffffffffffp-1,
                         12345.678901234567, -9876.54321, 0x1p52, 0x1.fffffffffffffp1023,
                         -0x1.8p1000, 0x1p-959, INFINITY, -INFINITY, NAN};
    const double ps[] = {1.0, 3.0, 0.1, 1.5, 0.7, 700.0, 0x1p-3, 0x1.fffffffffffffp-1,
                         3.7100000000000004, 0x1p-1001, 0x1.8p-1000, 0x1.8p1000, 0x1p1010,
                         0x1p-1070, -2.5};
    long applied = 0;
    for (double t : ts)
        for (double P : ps) {
            const double rP = gtls::fold_reciprocal(P);
            if (gtls::fold_time_small(t) || rP != rP) continue;
            ++applied;
            const double x = gtls::fold_quotient(t, P, rP), xd = t / P;
            const double ph = x - trunc(x), phd = xd - trunc(xd);
            if (!(same(ph, phd) || (ph != ph && phd != phd))) {
                ++bad_special;
                printf("special mismatch t=%a P=%a\n", t, P);
            }
        }
    // the small-time test on its boundaries
    const bool small_ok = !gtls::fold_time_small(0.0) && !gtls::fold_time_small(-0.0) &&
                          !gtls::fold_time_small(0x1p-960) && !gtls::fold_time_small(-0x1p-960) &&
                          gtls::fold_time_small(nextafter(0x1p-960, 0.0)) &&
                          gtls::fold_time_small(-0x1p-1074) && !gtls::fold_time_small(1.0);
    if (!small_ok || applied < 150) ++bad_special;
    printf("random %ld near_midpoint %ld cadence %ld special %ld naive_product_wrong %ld\n",
           bad_random, bad_mid, bad_cadence, bad_special, naive_wrong);
    return (bad_random || bad_mid || bad_cadence || bad_special || naive_wrong == 0) ? 1 : 0;
}
