/*
 * gtls_amd -- C ABI of the MI355X-native TLS chi2-vs-period search.
 *
 * Every array argument is a DEVICE pointer (memory the caller owns, e.g. a torch tensor
 * on the same device); sizes and scalars are passed by value. Nothing in this ABI
 * allocates device memory or synchronises: all work is enqueued on the context's
 * stream, into the caller's workspace, so a call can be captured into a HIP graph.
 *
 * Return value: 0 on success, GTLS_E_* (negative) on a library error, or
 * -(hipError_t) when a HIP call fails. gtls_last_error() gives a message.
 *
 * Reference interfaces these entry points replace (paths into tomarriott/GTLS):
 *   gtls_search / gtls_search_batched
 *       the per-batch kernel chain of src/gputls/core.py:691-889 (search_multi_periods):
 *       foldFast (src/gputls/GPUFun.cu:28-39), CuPy argsort (core.py:775-777),
 *       patchData (GPUFun.cu:108-127), calcInverseSquaredPatchedDy (GPUFun.cu:132-140),
 *       calcEdgeEffectCorrections (GPUFun.cu:145-166), cp.cumsum (core.py:799-822),
 *       calculate_base_error / calcAllFullSum_v2 / calculate_final_ootr_v3
 *       (GPUFun.cu:205-357), calcAllLowestResidualsGPUB_SignalTiled_v2
 *       (GPUFun.cu:380-436) and cp.argmin (core.py:866-875); the no-skip variants of the
 *       re-search (core.py:1001-1247, GPUFun.cu:817-871) are the same call with
 *       skip_point = GTLS_SKIP_NONE. The batched form runs many light curves, each with
 *       its own period grid and width table, in one call.
 *   gtls_fold_sort
 *       foldFast + argsort (GPUFun.cu:28-39, core.py:775-777): the stable phase-order
 *       permutation of every trial period, exposed for bit-exact checks.
 *   gtls_spectra_batched
 *       spectra() of src/gputls/stats.py:95-122 (SR, power_raw, detrended power, SDE)
 *       with the running median of src/gputls/helpers.py:91-106, per light curve.
 *   gtls_transit_fwd / gtls_transit_bwd / gtls_transit_chi2 / gtls_transit_chi2_grouped
 *       no reference counterpart: the model of the gradient refinement of grid peaks.
 *   gtls_set_timing / gtls_kernel_ms / gtls_set_tuning
 *       no reference counterpart: per-kernel timing and tuning of the search chain.
 */
#ifndef GTLS_AMD_H
#define GTLS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GTLS_OK 0
#define GTLS_E_SHAPE (-1001)     /* inconsistent sizes or a null pointer */
#define GTLS_E_WORKSPACE (-1002) /* workspace smaller than gtls_*_workspace_bytes() */
#define GTLS_E_MODE (-1003)      /* unknown arithmetic mode */
#define GTLS_E_DEVICE (-1004)    /* device not usable / not gfx950 */

#define GTLS_MODE_FP64 0  /* float64 throughout */
#define GTLS_MODE_FP32C 1 /* float64 fold, prefix sums and final chi2; float32 windows */

#define GTLS_SKIP_NONE 0x7f800000LL /* skip_point value for T0_fit_margin == 0 */

typedef struct gtls_ctx gtls_ctx;

/*
 * A batch of light curves in CSR layout. Curve c owns points [pt_off[c], pt_off[c+1]),
 * periods [per_off[c], per_off[c+1]) and widths [wid_off[c], wid_off[c+1]).
 * For period g of curve c the searched widths are the curve-local indices
 * k_lo[g]..k_hi[g] (inclusive; empty if k_lo > k_hi) of the curve's ascending width
 * table. Width k (global index wid_off[c] + k) has template row
 * tmpl[tmpl_off[.] .. tmpl_off[.] + widths[.]) holding S = 1 - signal and overshoot
 * factor overshoot[.]. Pointers are device pointers; the size fields are host values.
 */
typedef struct gtls_batch {
    int32_t n_curves;
    const int64_t* pt_off;   /* [n_curves + 1] */
    const double* t;         /* [total_points] time, days */
    const double* y;         /* [total_points] normalised flux */
    const double* dy;        /* [total_points] flux error */
    const int64_t* per_off;  /* [n_curves + 1] */
    const double* periods;   /* [total_periods] days */
    const int32_t* k_lo;     /* [total_periods] */
    const int32_t* k_hi;     /* [total_periods] */
    const int64_t* wid_off;  /* [n_curves + 1] */
    const int32_t* widths;   /* [total_widths] samples, ascending within a curve */
    const double* overshoot; /* [total_widths] */
    const int64_t* tmpl_off; /* [total_widths] */
    const double* tmpl;      /* [total_tmpl] */
    int64_t skip_point;      /* int(1 / T0_fit_margin), or GTLS_SKIP_NONE */
    double depth_min;        /* transit_depth_min */
    int32_t negative_time;   /* nonzero if any t < 0 (changes the bucket map only) */
    /* host-side sizes (must match the device arrays) */
    int64_t total_points;
    int64_t total_periods;
    int64_t total_widths;
    int64_t total_tmpl;
    int32_t max_points;      /* max over curves of the point count */
    int32_t max_width;       /* max over all widths */
    int32_t max_period_widths; /* max over periods of k_hi - k_lo + 1 (an upper bound is
                                  fine, e.g. the largest width table of a curve) */
} gtls_batch;

/* Context bound to one device and one stream (hipStream_t, may be NULL). */
int gtls_create(int device, void* stream, gtls_ctx** out);
void gtls_destroy(gtls_ctx* ctx);
int gtls_set_stream(gtls_ctx* ctx, void* stream);
const char* gtls_last_error(const gtls_ctx* ctx);
/* Persistent workgroups the search launches on this device (one scratch slot each). */
int32_t gtls_search_slots(const gtls_ctx* ctx);

/* Workspace bytes gtls_search_batched needs for this batch and mode (>= 0), or < 0. */
int64_t gtls_search_workspace_bytes(gtls_ctx* ctx, int32_t mode, const gtls_batch* batch);

/*
 * chi2-vs-period search of every period of every curve of the batch.
 * Outputs per period g (device arrays of length total_periods): chi2[g] = minimum of the
 * TLS statistic over (width, T0 position) (NaN if the period has no widths),
 * best_k[g] = curve-local width index of that minimum, best_s[g] = its position in the
 * phase-sorted sequence; ties resolve to the first in (width, position) order.
 */
int gtls_search_batched(gtls_ctx* ctx, int32_t mode, const gtls_batch* batch,
                        void* workspace, int64_t workspace_bytes,
                        double* chi2, int32_t* best_k, int32_t* best_s);

/* Single light curve: the same search with the arrays of one curve (device pointers). */
int64_t gtls_search_workspace_bytes_1(gtls_ctx* ctx, int32_t mode, int64_t n,
                                      int64_t n_periods, int32_t n_widths,
                                      int64_t total_tmpl, int32_t max_width);
int gtls_search(gtls_ctx* ctx, int32_t mode,
                const double* t, const double* y, const double* dy, int64_t n,
                const double* periods, int64_t n_periods,
                const int32_t* k_lo, const int32_t* k_hi,
                const int32_t* widths, int32_t n_widths, int32_t max_width,
                const double* overshoot, const int64_t* tmpl_off, const double* tmpl,
                int64_t total_tmpl, int64_t skip_point, double depth_min,
                int32_t negative_time, void* workspace, int64_t workspace_bytes,
                double* chi2, int32_t* best_k, int32_t* best_s);

/* Phase-order permutation of t for each period: perm[p * n + j] (stable in index). */
int64_t gtls_fold_sort_workspace_bytes(gtls_ctx* ctx, int64_t n);
int gtls_fold_sort(gtls_ctx* ctx, const double* t, int64_t n, const double* periods,
                   int64_t n_periods, int32_t negative_time,
                   void* workspace, int64_t workspace_bytes, int32_t* perm);

/*
 * SR / SDE spectra of every curve of a batch from its chi2 vector (CSR by per_off):
 * power_raw, power (detrended with a running median of odd width `kernel`, applied when a
 * curve has more than 2 * kernel periods), SDE_raw and SDE per curve. status[c] = 0 when
 * the device computed curve c, 1 when the curve needs the host path (a NaN chi2, or
 * max chi2 > 100 * min chi2, the only case in which the outlier mask can be non-empty).
 * Workspace: gtls_spectra_workspace_bytes(total_periods).
 */
int64_t gtls_spectra_workspace_bytes(gtls_ctx* ctx, int64_t total_periods);
int gtls_spectra_batched(gtls_ctx* ctx, int32_t n_curves, const int64_t* per_off,
                         const double* chi2, int64_t total_periods, int32_t kernel,
                         void* workspace, int64_t workspace_bytes, double* power_raw,
                         double* power, double* sde_raw, double* sde, int32_t* status);

/*
 * gtls_spectra_batched without a workspace, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * No counterpart in the reference. The running-median buffer is a file-scope device array
 * of the library, 4194304 float64 (32 MiB) per GPU, held until the process ends; calls
 * with more periods return GTLS_E_WORKSPACE. The call allocates nothing. Calls that
 * overlap on one GPU (concurrent streams or threads) share the buffer and race; calls in
 * one stream order do not.
 */
int gtls_spectra_batched_noalloc(gtls_ctx* ctx, int32_t n_curves, const int64_t* per_off,
                                 const double* chi2, int64_t total_periods, int32_t kernel,
                                 double* power_raw, double* power, double* sde_raw,
                                 double* sde, int32_t* status);

/*
 * Kernel timing of the search (no reference counterpart; for measuring the chain's kernels
 * from the calling program). With gtls_set_timing(ctx, 1) every later search call records
 * HIP events on the context's stream around each period chunk's prep, windows and reduce
 * kernels (the context keeps and reuses the events; gtls_set_timing also discards earlier
 * records). gtls_kernel_ms waits for the last recorded event, writes the summed
 * milliseconds of the prep, windows and reduce kernels since the previous call into
 * ms[0..2], clears the records and returns the number of chunks (or a negative error).
 * No events are recorded while the stream is being captured into a graph (see
 * gtls_set_graph_timing); more than 16384 period chunks between two gtls_kernel_ms calls
 * discard the records, and that call returns GTLS_E_SHAPE.
 */
int gtls_set_timing(gtls_ctx* ctx, int32_t on);

/*
 * Kernel timing that survives graph capture (no reference counterpart). With a device
 * buffer `stamps` of 4 * max_chunks uint64 (caller-owned, max_chunks <= 16384), every later
 * search call on the context -- run now or captured into a graph -- launches a one-thread
 * kernel before and after each period chunk's prep, windows and reduce kernels that
 * stores the device wall clock into the next slot; every replay of a captured graph writes
 * the same slots again. gtls_kernel_ms then waits for the device and returns the figures
 * of the last run or replay (without clearing them); stamps = NULL switches this off.
 * Use one context per timed graph.
 */
int gtls_set_graph_timing(gtls_ctx* ctx, uint64_t* stamps, int64_t max_chunks);

/*
 * Tuning of the float32 search (no reference counterpart; set by gtls_amd.tune): strides q
 * below toep_q (0 .. 24, default 8) take the Toeplitz matrix-core tiles, larger ones the
 * diagonal-sum tiles; chunk_mb (0 keeps the current value, below 2^43) bounds the
 * per-period buffers of one period chunk; wg_per_cu (0: as many as fit) caps the resident workgroups per CU
 * of the persistent search kernels (their waves per SIMD); with a common weight, the
 * diagonal-sum tiles of strides q >= quad_q (0: the default, 96) load their sample
 * operands 8 rows x 128 bytes per load and move them to the matrix-core layout in
 * registers, shorter strides load them in that layout directly (the results are the same
 * either way; with per-sample weights the bound is fixed at 24). Applies to later calls on
 * the context; an argument out of range returns GTLS_E_SHAPE and changes nothing.
 */
int gtls_set_tuning(gtls_ctx* ctx, int32_t toep_q, int64_t chunk_mb, int32_t wg_per_cu,
                    int32_t quad_q);
int64_t gtls_kernel_ms(gtls_ctx* ctx, double* ms);

/*
 * Smooth limb-darkened transit model for refining grid peaks by gradient descent. No
 * counterpart in the reference (it refines on the host, src/gputls/stats.py:233-488,
 * without a model fit); these two calls are the forward and backward of one
 * torch.autograd.Function.
 *
 * Model m (theta[6 m .. 6 m + 5] = period, T0, duration, depth, q1, q2; float64) covers
 * the points [pt_off[c], pt_off[c + 1]) of curve c = model_curve[m] (times t, float64)
 * and owns the flux slots [out_off[m], out_off[m] + N_c). Fluxes and their gradients are
 * float64 in GTLS_MODE_FP64 and float32 in GTLS_MODE_FP32C (the fold is float64 in both).
 *
 * gtls_transit_fwd writes the model flux of every point; max_points (the largest N_c) only
 * sizes the launch; n_models <= 65535.
 * gtls_transit_bwd writes grad_theta[6 m + k] = sum_i grad_flux[out_off[m] + i] *
 * d flux_i / d theta_k, accumulated in float64.
 * gtls_transit_chi2 fuses the model with the fit's loss and needs no flux slots: with the
 * curves' fluxes y and weights w (float64, per point like t) it writes
 * chi2[m] = sum_i w_i (y_i - f_i)^2 and, unless grad_theta is NULL,
 * grad_theta[6 m + k] = d chi2[m] / d theta_k, in float64.
 */
int gtls_transit_fwd(gtls_ctx* ctx, int32_t mode, int64_t n_models, const double* theta,
                     const int32_t* model_curve, const int64_t* pt_off, const double* t,
                     const int64_t* out_off, int64_t max_points, void* flux);
int gtls_transit_bwd(gtls_ctx* ctx, int32_t mode, int64_t n_models, const double* theta,
                     const int32_t* model_curve, const int64_t* pt_off, const double* t,
                     const int64_t* out_off, const void* grad_flux, double* grad_theta);
int gtls_transit_chi2(gtls_ctx* ctx, int32_t mode, int64_t n_models, const double* theta,
                      const int32_t* model_curve, const int64_t* pt_off, const double* t,
                      const double* y, const double* w, double* chi2, double* grad_theta);

/*
 * gtls_transit_chi2 with the models grouped by curve (no reference counterpart either):
 * the models of curve c are [model_off[c], model_off[c + 1]) (at most max_models of them),
 * theta in that order. The points of curve c are cut into tiles of
 * gtls_transit_tile_points() consecutive points, tiles [tile_off[c], tile_off[c + 1]) with
 * tile_off[c + 1] - tile_off[c] = ceil(N_c / tile points) and n_tiles = tile_off[n_curves];
 * each tile is read once and every model of its curve is evaluated on it, so a call reads
 * each point once whatever the number of models; a point farther than 0.71 durations from
 * a model's transit (outside the model: flux 1) adds the flat model's w (y - 1)^2, summed
 * once per tile for all models. part is caller scratch of part_len >=
 * n_tiles * (max_models * (grad_theta ? 7 : 1) + 1) doubles for the per-tile sums, which
 * are added in tile order (results do not depend on the launch). Outputs as
 * gtls_transit_chi2 (the same sums, in another order).
 */
int64_t gtls_transit_tile_points(void);
int gtls_transit_chi2_grouped(gtls_ctx* ctx, int32_t mode, int32_t n_curves, int64_t n_models,
                              int32_t max_models, const double* theta, const int64_t* model_off,
                              const int64_t* pt_off, const int64_t* tile_off, int64_t n_tiles,
                              const double* t, const double* y, const double* w, double* part,
                              int64_t part_len, double* chi2, double* grad_theta);

#ifdef __cplusplus
}
#endif
#endif /* GTLS_AMD_H */
