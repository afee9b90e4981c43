/*
 * FP64 CPU oracle of the TLS chi2-vs-period search -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with OpenMP over periods, the arithmetic of one trial
 * period of the reference search, in the reference's own structure (prefix sums of the
 * patched flux and of the weighted squared residual, an out-of-transit residual looked
 * up from the prefix sums, a direct in-transit residual, an edge correction):
 *
 *   fold            src/gputls/GPUFun.cu:28-39   phase = x - trunc(x), x = t/P
 *   argsort         src/gputls/core.py:775-777   stable order of the phases
 *   patch/gather    src/gputls/GPUFun.cu:108-127 Y_i = y[perm(i mod N)], i < N + W
 *   weights         src/gputls/GPUFun.cu:132-140 w = 1/dy^2
 *   edge correction src/gputls/GPUFun.cu:145-166 sum over i in [N, N+W) of (1-Y)^2 w
 *   prefix sums     src/gputls/core.py:799-822   C = cumsum(Y), E = cumsum((1-Y)^2 w)
 *   out-of-transit  src/gputls/GPUFun.cu:205-230,315-357
 *   window mean     src/gputls/GPUFun.cu:364-373
 *   in-transit stat src/gputls/GPUFun.cu:380-436 (gate, skip stride, sentinel = N)
 *   argmin          src/gputls/core.py:866-875   first minimum, width-major, T0-minor
 *
 * The per-period width set is the inclusive index range [k_lo[p], k_hi[p]] into the
 * ascending width table (the reference evaluates a batch union instead; the engine
 * defines the set per period).
 *
 * It is the checker of the HIP search and its CPU baseline; the product never calls it.
 * Its own pins live in tests/test_oracle.py (direct-sum brute force on small inputs,
 * closed-form special cases).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double key;
    int32_t idx;
} keyed_t;

static int cmp_keyed(const void* a, const void* b) {
    const keyed_t* x = (const keyed_t*)a;
    const keyed_t* y = (const keyed_t*)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

/* Phase of one sample: x - trunc(x) with x = t / P (src/gputls/GPUFun.cu:36-37). */
static inline double fold_phase(double t, double P) {
    double x = t / P;
    return x - trunc(x);
}

/* Stable argsort of the folded phases of t for period P. */
void tls_oracle_fold_sort(const double* t, int64_t n, double P, int32_t* perm_out) {
    keyed_t* buf = (keyed_t*)malloc(sizeof(keyed_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t j = 0; j < n; ++j) {
        buf[j].key = fold_phase(t[j], P);
        buf[j].idx = (int32_t)j;
    }
    qsort(buf, (size_t)n, sizeof(keyed_t), cmp_keyed);
    for (int64_t j = 0; j < n; ++j) perm_out[j] = buf[j].idx;
    free(buf);
}

typedef struct {
    keyed_t* keyed;
    double* Y;   /* patched flux, length n + wmax */
    double* w;   /* patched weights */
    double* Cs;  /* prefix of Y */
    double* Es;  /* prefix of (1-Y)^2 w */
} work_t;

static int work_alloc(work_t* wk, int64_t n, int64_t np_) {
    wk->keyed = (keyed_t*)malloc(sizeof(keyed_t) * (size_t)n);
    wk->Y = (double*)malloc(sizeof(double) * (size_t)np_);
    wk->w = (double*)malloc(sizeof(double) * (size_t)np_);
    wk->Cs = (double*)malloc(sizeof(double) * (size_t)np_);
    wk->Es = (double*)malloc(sizeof(double) * (size_t)np_);
    return wk->keyed && wk->Y && wk->w && wk->Cs && wk->Es;
}

static void work_free(work_t* wk) {
    free(wk->keyed); free(wk->Y); free(wk->w); free(wk->Cs); free(wk->Es);
}

/*
 * One period. stat_out, when not NULL, receives the full statistic for every
 * (width index in range, position s < n) -- row (k - k_lo), column s.
 */
static void search_one(const double* t, const double* y, const double* dy, int64_t n,
                       double P, const int32_t* widths, int32_t wmax,
                       const double* templ, const int64_t* templ_off,
                       const double* overshoot, int32_t k_lo, int32_t k_hi,
                       int64_t skip_point, double depth_min, work_t* wk,
                       double* chi2, int32_t* best_k, int32_t* best_s, double* stat_out) {
    const int64_t np_ = n + wmax;
    for (int64_t j = 0; j < n; ++j) {
        wk->keyed[j].key = fold_phase(t[j], P);
        wk->keyed[j].idx = (int32_t)j;
    }
    qsort(wk->keyed, (size_t)n, sizeof(keyed_t), cmp_keyed);

    for (int64_t i = 0; i < np_; ++i) {
        int32_t src = wk->keyed[i % n].idx;
        double d = dy[src];
        wk->Y[i] = y[src];
        wk->w[i] = 1.0 / (d * d);
    }
    double edge = 0.0;
    for (int64_t i = n; i < np_; ++i) {
        double r = 1.0 - wk->Y[i];
        edge += r * r * wk->w[i];
    }
    double cacc = 0.0, eacc = 0.0;
    for (int64_t i = 0; i < np_; ++i) {
        double r = 1.0 - wk->Y[i];
        cacc += wk->Y[i];
        eacc += r * r * wk->w[i];
        wk->Cs[i] = cacc;
        wk->Es[i] = eacc;
    }
    const double e_total = wk->Es[np_ - 1];
    const double sentinel = (double)n;

    double best = INFINITY;
    int32_t bk = -1, bs = -1;
    for (int32_t k = k_lo; k <= k_hi; ++k) {
        const int64_t d = widths[k];
        if (d <= 0) {
            /* empty window: its mean is 0/0, so every position carries the sentinel */
            for (int64_t s = 0; s < n; ++s) {
                if (stat_out) stat_out[(int64_t)(k - k_lo) * n + s] = sentinel;
                if (sentinel < best) { best = sentinel; bk = k; bs = (int32_t)s; }
            }
            continue;
        }
        const int64_t skip = (d > skip_point) ? (d / skip_point) : 1;
        const double* S = templ + templ_off[k];
        const double os = overshoot[k];
        const double fullsum = e_total - wk->Es[d - 1];
        for (int64_t s = 0; s < n; ++s) {
            double mean = (s == 0) ? 1.0 - wk->Cs[d - 1] / (double)d
                                   : 1.0 - (wk->Cs[s + d - 1] - wk->Cs[s - 1]) / (double)d;
            double stat = sentinel;
            if (mean > depth_min && (s % skip) == 0) {
                double ootr = (s == 0) ? fullsum
                                       : fullsum + (wk->Es[s - 1] - (wk->Es[s - 1 + d] - wk->Es[d - 1]));
                double scale = mean * os * 2.0;
                double res = 0.0;
                for (int64_t i = 0; i < d; ++i) {
                    double loss = wk->Y[s + i] - (1.0 - S[i] * scale);
                    res += loss * loss * wk->w[s + i];
                }
                stat = res + ootr - edge;
            }
            if (stat_out) stat_out[(int64_t)(k - k_lo) * n + s] = stat;
            if (stat < best) {
                best = stat;
                bk = k;
                bs = (int32_t)s;
            }
        }
    }
    if (bk < 0) {
        *chi2 = NAN;
        *best_k = -1;
        *best_s = -1;
    } else {
        *chi2 = best;
        *best_k = bk;
        *best_s = bs;
    }
}

/* Full search over all periods; returns 0, or -1 on allocation failure. */
int tls_oracle_search(const double* t, const double* y, const double* dy, int64_t n,
                      const double* periods, int64_t n_periods,
                      const int32_t* widths, int32_t n_widths, int32_t wmax,
                      const double* templ, const int64_t* templ_off, const double* overshoot,
                      const int32_t* k_lo, const int32_t* k_hi,
                      int64_t skip_point, double depth_min, int32_t nthreads,
                      double* chi2_out, int32_t* best_k_out, int32_t* best_s_out) {
    (void)n_widths;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        work_t wk;
        int ok = work_alloc(&wk, n, n + wmax);
        if (!ok) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 4)
        for (int64_t p = 0; p < n_periods; ++p) {
            if (!ok) continue;
            search_one(t, y, dy, n, periods[p], widths, wmax, templ, templ_off, overshoot,
                       k_lo[p], k_hi[p], skip_point, depth_min, &wk,
                       &chi2_out[p], &best_k_out[p], &best_s_out[p], NULL);
        }
        work_free(&wk);
    }
    return failed ? -1 : 0;
}

/* One period with the whole statistic table (rows k_lo..k_hi, n columns). */
int tls_oracle_period_table(const double* t, const double* y, const double* dy, int64_t n,
                            double P, const int32_t* widths, int32_t wmax,
                            const double* templ, const int64_t* templ_off,
                            const double* overshoot, int32_t k_lo, int32_t k_hi,
                            int64_t skip_point, double depth_min, double* stat_out,
                            double* chi2, int32_t* best_k, int32_t* best_s) {
    work_t wk;
    if (!work_alloc(&wk, n, n + wmax)) {
        work_free(&wk);
        return -1;
    }
    search_one(t, y, dy, n, P, widths, wmax, templ, templ_off, overshoot, k_lo, k_hi,
               skip_point, depth_min, &wk, chi2, best_k, best_s, stat_out);
    work_free(&wk);
    return 0;
}
