// C ABI of the TLS search (declared in include/gtls_amd.h): argument checks, workspace
// layout and kernel launches. Everything is enqueued on the context's stream; nothing here
// allocates device memory or waits for the device.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gtls_amd.h"
#include "search.h"
#include "spectra.h"
#include "transit_model.h"

struct gtls_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    int prep_fp64 = 0, prep_fp32 = 0;   // resident workgroups of prep_kernel
    int win_fp64 = 0, win_fp32 = 0;     // resident workgroups of windows_kernel
    int slots_sort = 0;
    int use_mfma = 1;
    int toep_q = gtls::kToepQDefault;
    int quad_q = gtls::kQuadQ;          // common-weight strides from this on: Quad operand loads
    int wg_per_cu = 0;                  // cap on resident search workgroups per CU (0: none)
    int64_t chunk_bytes = 0;            // budget of one chunk's gathered samples
    std::string err;
    // kernel timing (gtls_set_timing): per period chunk, events before the prep kernel,
    // between prep and windows, between windows and reduce, after reduce
    int timing = 0;
    std::vector<hipEvent_t> pool;       // events owned by the context
    size_t used = 0;                    // events recorded since the last gtls_kernel_ms
    bool overflow = false;              // more records than kMaxTimingEvents were asked for
    double* spectra_scratch = nullptr;  // this device's copy of `scratch` (device address)
    // graph timing (gtls_set_graph_timing): caller-owned device buffer of wall-clock stamps
    uint64_t* stamps = nullptr;
    int64_t stamp_cap = 0;              // stamps the buffer holds
    double wall_khz = 0.0;              // wall-clock counter rate
    ~gtls_ctx() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

namespace {

using namespace gtls;

constexpr int64_t kAlign = 256;
constexpr size_t kMaxTimingEvents = 1 << 16;   // records kept between gtls_kernel_ms calls

// The device wall clock (a constant-rate counter) at the time this kernel runs.
__global__ void wall_stamp_kernel(uint64_t* out) { *out = wall_clock64(); }

int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

int fail(gtls_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(gtls_ctx* ctx, hipError_t e, const char* where) {
    if (ctx) ctx->err = std::string(where) + ": " + hipGetErrorString(e);
    return -(int)e;
}

struct SearchLayout {
    int64_t uw0 = 0, tmpl2 = 0, tmpl1 = 0, tmplsq = 0, ss = 0, wconst = 0, flat = 0, counters = 0;
    int64_t tmpl_tail = 0;
    int64_t tmplh = 0, tmplc = 0, xsc = 0, upair = 0, n_chunked = 0, n_zero = 0, tplen = 0;
    int64_t cplen = 0;
    int64_t slots = 0, slot_bytes = 0, o_key = 0, o_perm = 0, o_tmp_key = 0, o_tmp_id = 0;
    int64_t o_lists = 0, n_slots = 0;
    int64_t xbuf = 0, dbuf = 0, meta = 0, part = 0;
    int64_t xstride = 0, chunk = 0, n_chunks = 0, kmax = 0, n_win = 0;
    int64_t total = 0;
};

// Workspace: per-launch prepared copies, per-workgroup sort scratch, and the buffers of one
// chunk of periods (gathered samples, running sums, per-period meta, per-width minima).
SearchLayout plan_search(const gtls_ctx* ctx, int mode, const gtls_batch* b) {
    const int64_t real = (mode == GTLS_MODE_FP64) ? 8 : 4;
    const int64_t max_points = b->max_points, max_width = b->max_width > 0 ? b->max_width : 1;
    SearchLayout L;
    int64_t off = 0;
    L.uw0 = off;
    off = align_up(off + 2 * real * b->total_points);
    // rows with kTmplPad zeros on both sides, then a zeroed tail of at least one widest
    // row past the last row (over-reads of the window loops)
    L.tmpl_tail = std::max<int64_t>(kTmplTail, align_up(max_width + 1024));
    const int64_t tmpl_padded = b->total_tmpl + 2 * kTmplPad * b->total_widths + L.tmpl_tail;
    L.tmpl2 = off;
    off = align_up(off + 2 * real * tmpl_padded);
    L.tmpl1 = off;
    off = align_up(off + real * tmpl_padded);
    L.tmplsq = off;
    off = align_up(off + real * tmpl_padded);
    // float32: float16 pair planes of the template rows (Toeplitz tiles) and of the chunked
    // rows with their zero blocks (diagonal-sum tiles), two planes each (S and S^2)
    if (mode == GTLS_MODE_FP32C) {
        L.tplen = tmpl_padded;
        L.n_chunked = 4 * b->total_tmpl + 1024 * b->total_widths + 1024;
        L.n_zero = max_width + 128;
        L.cplen = L.n_chunked + L.n_zero;
    }
    L.tmplh = off;
    off = align_up(off + 8 * L.tplen);
    L.tmplc = off;
    off = align_up(off + 8 * L.cplen);
    L.xsc = off;
    off = align_up(off + 8 * (int64_t)b->n_curves);
    L.upair = off;
    off = align_up(off + (mode == GTLS_MODE_FP32C ? 4 * b->total_points : 0));
    L.ss = off;
    off = align_up(off + 8 * (b->total_widths > 0 ? b->total_widths : 1));
    L.wconst = off;
    off = align_up(off + 8 * (int64_t)b->n_curves);
    L.flat = off;
    off = align_up(off + 8 * (int64_t)b->n_curves);
    // sort scratch of one prep workgroup
    int64_t s = 0;
    L.o_key = s;
    s = align_up(s + 8 * max_points);
    L.o_perm = s;
    s = align_up(s + 4 * max_points);
    L.o_tmp_key = s;
    s = align_up(s + 8 * max_points);
    L.o_tmp_id = s;
    s = align_up(s + 4 * max_points);
    L.o_lists = s;
    s = align_up(s + 8 * (int64_t)kMaxBuckets);
    L.slot_bytes = s;
    L.n_slots = (mode == GTLS_MODE_FP64) ? ctx->prep_fp64 : ctx->prep_fp32;
    L.n_win = (mode == GTLS_MODE_FP64) ? ctx->win_fp64 : ctx->win_fp32;
    if (ctx->wg_per_cu > 0) {
        L.n_slots = std::min<int64_t>(L.n_slots, (int64_t)ctx->n_cu * ctx->wg_per_cu);
        L.n_win = std::min<int64_t>(L.n_win, (int64_t)ctx->n_cu * ctx->wg_per_cu);
    }
    L.slots = off;
    off = align_up(off + L.slot_bytes * L.n_slots);
    // per period of a chunk: gathered samples N + widest width of the wrap, plus the zeros
    // the window loops read past them (up to another max_width + 96, or 416)
    L.xstride = max_points + 2 * max_width + 512;
    L.kmax = b->max_period_widths;
    const int64_t per_period =
        L.xstride * (2 * real + 8) + (int64_t)sizeof(PeriodMeta) + L.kmax * (int64_t)sizeof(ArgMin);
    const int64_t periods = b->total_periods > 0 ? b->total_periods : 1;
    L.chunk = std::max<int64_t>(1, std::min<int64_t>(periods, ctx->chunk_bytes / per_period));
    L.n_chunks = (periods + L.chunk - 1) / L.chunk;
    L.counters = off;
    off = align_up(off + kChunkCounters * 8 * L.n_chunks);
    L.xbuf = off;
    off = align_up(off + 2 * real * L.xstride * L.chunk);
    L.dbuf = off;
    off = align_up(off + 8 * L.xstride * L.chunk);
    L.meta = off;
    off = align_up(off + (int64_t)sizeof(PeriodMeta) * L.chunk);
    L.part = off;
    off = align_up(off + (int64_t)sizeof(ArgMin) * L.kmax * L.chunk);
    L.total = align_up(off);
    return L;
}

bool batch_ok(const gtls_batch* b, std::string& why) {
    if (!b) { why = "batch is NULL"; return false; }
    if (b->n_curves <= 0) { why = "n_curves must be positive"; return false; }
    if (!b->pt_off || !b->t || !b->y || !b->dy || !b->per_off || !b->periods || !b->k_lo ||
        !b->k_hi || !b->wid_off || !b->widths || !b->overshoot || !b->tmpl_off || !b->tmpl) {
        why = "a batch array pointer is NULL";
        return false;
    }
    if (b->total_points <= 0 || b->total_periods < 0 || b->total_widths <= 0 ||
        b->total_tmpl < 0 || b->max_points <= 0 || b->max_width < 0) {
        why = "batch sizes out of range";
        return false;
    }
    if (b->max_points > (1LL << 30) || b->max_width > (1LL << 30)) {
        why = "a light curve or width exceeds 2^30 samples";
        return false;
    }
    if (b->skip_point <= 0) { why = "skip_point must be positive"; return false; }
    if (b->max_period_widths <= 0 || b->max_period_widths > (1 << 20)) {
        why = "max_period_widths out of range";
        return false;
    }
    return true;
}

// The float32 tiles address planes with 32-bit byte offsets.
bool layout_ok(const SearchLayout& L, int mode, std::string& why) {
    if (mode != GTLS_MODE_FP32C) return true;
    if (4 * L.tplen >= (1LL << 30) || 4 * L.cplen >= (1LL << 30) || 8 * L.xstride >= (1LL << 32)) {
        why = "batch too large for the float32 search (template planes over 2^30 bytes or a "
              "light curve over 2^29 samples): split it";
        return false;
    }
    return true;
}

__global__ void zero_u64s_kernel(unsigned long long* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ULL;
}

template <class Real>
int launch_search(gtls_ctx* ctx, const gtls_batch* b, char* ws, const SearchLayout& L,
                  double* chi2, int32_t* best_k, int32_t* best_s) {
    using R2 = typename Pair<Real>::type;
    hipStream_t st = ctx->stream;
    R2* uw0 = (R2*)(ws + L.uw0);
    R2* tmpl2 = (R2*)(ws + L.tmpl2);
    Real* tmpl1 = (Real*)(ws + L.tmpl1);
    Real* tmplsq = (Real*)(ws + L.tmplsq);
    double* ss = (double*)(ws + L.ss);
    double* wconst = (double*)(ws + L.wconst);
    double* flat = (double*)(ws + L.flat);
    unsigned long long* counters = (unsigned long long*)(ws + L.counters);

    const int tpb = 256;
    hipLaunchKernelGGL(prep_points_kernel<Real>, dim3((unsigned)((b->total_points + tpb - 1) / tpb)),
                       dim3(tpb), 0, st, b->y, b->dy, b->total_points, uw0);
    hipLaunchKernelGGL(prep_template_kernel<Real>, dim3((unsigned)b->total_widths + 8), dim3(tpb),
                       0, st, b->tmpl, b->tmpl_off, b->widths, b->total_widths, b->total_tmpl,
                       L.tmpl_tail, tmpl2, tmpl1, tmplsq,
                       L.tplen > 0 ? (uint32_t*)(ws + L.tmplh) : (uint32_t*)nullptr, L.tplen);
    if (L.n_chunked > 0)
        hipLaunchKernelGGL(chunk_template_kernel, dim3((unsigned)b->total_widths + 4), dim3(tpb), 0,
                           st, b->tmpl, b->tmpl_off, b->widths, b->total_widths, b->skip_point,
                           (uint32_t*)(ws + L.tmplc), L.cplen, L.n_chunked, L.n_zero);
    hipLaunchKernelGGL(template_ss_kernel<Real>, dim3((unsigned)b->total_widths), dim3(kWave), 0,
                       st, b->tmpl, b->tmpl_off, b->widths, b->total_widths, ss);
    hipLaunchKernelGGL(chi2_flat_kernel, dim3(b->n_curves), dim3(kBlock), 0, st, b->pt_off,
                       b->y, b->dy, flat, wconst, (float2*)(ws + L.xsc));
    if (sizeof(Real) == 4)
        hipLaunchKernelGGL(pair_points_kernel, dim3(b->n_curves), dim3(kBlock), 0, st, b->pt_off,
                           (const float2*)uw0, (const float2*)(ws + L.xsc),
                           (uint32_t*)(ws + L.upair));
    const int64_t ncount = (int64_t)kChunkCounters * L.n_chunks;
    hipLaunchKernelGGL(zero_u64s_kernel, dim3((unsigned)std::min<int64_t>(1024, (ncount + tpb - 1) / tpb)),
                       dim3(tpb), 0, st, counters, ncount);

    SearchParams<Real> p;
    p.n_curves = b->n_curves;
    p.pt_off = b->pt_off;
    p.t = b->t;
    p.per_off = b->per_off;
    p.periods = b->periods;
    p.k_lo = b->k_lo;
    p.k_hi = b->k_hi;
    p.wid_off = b->wid_off;
    p.widths = b->widths;
    p.overshoot = b->overshoot;
    p.tmpl_off = b->tmpl_off;
    p.uw0 = uw0;
    p.tmpl2 = tmpl2;
    p.tmpl1 = tmpl1;
    p.tmplsq = tmplsq;
    p.tmplh = ws + L.tmplh;
    p.tmplc = ws + L.tmplc;
    p.tlo = (int32_t)(4 * L.tplen);
    p.clo = (int32_t)(4 * L.cplen);
    p.zoff = L.n_chunked;
    p.xsc = (const float2*)(ws + L.xsc);
    p.upair = (const uint32_t*)(ws + L.upair);
    // curves with a common weight need the prefix sums D only if some width of the batch
    // takes a path that reads them (the tiles compute their window means themselves)
    p.d_constw = 0;
    if (sizeof(Real) == 4) {
        for (int32_t d = 1; d <= std::max<int32_t>(1, b->max_width); ++d) {
            const int32_t q = (int32_t)((d > b->skip_point) ? (d / b->skip_point) : 1);
            const bool toep = ctx->use_mfma && q < ctx->toep_q;
            const bool diag = ctx->use_mfma && !toep && q >= 4 && d / q + 1 < kOnesCol;
            if (!toep && !diag) p.d_constw = 1;
        }
    }
    p.ss = ss;
    p.chi2_flat = flat;
    p.wconst = wconst;
    p.skip_point = b->skip_point;
    p.depth_min = b->depth_min;
    p.negative_time = b->negative_time;
    p.use_mfma = ctx->use_mfma;
    p.toep_q = ctx->toep_q;
    p.quad_q = ctx->quad_q;
    p.max_points = b->max_points;
    p.max_width = b->max_width > 0 ? b->max_width : 1;
    p.kmax = (int32_t)L.kmax;
    p.slots = ws + L.slots;
    p.slot_bytes = L.slot_bytes;
    p.o_key = L.o_key;
    p.o_perm = L.o_perm;
    p.o_tmp_key = L.o_tmp_key;
    p.o_tmp_id = L.o_tmp_id;
    p.o_lists = L.o_lists;
    p.xbuf = (R2*)(ws + L.xbuf);
    p.dbuf = (double*)(ws + L.dbuf);
    p.xstride = L.xstride;
    p.meta = (PeriodMeta*)(ws + L.meta);
    p.part = (ArgMin*)(ws + L.part);
    p.chi2 = chi2;
    p.best_k = best_k;
    p.best_s = best_s;
    // With timing on, a mark before and after each of the chunk's three kernels. Mode 1:
    // HIP events, for calls that run now (nothing while the stream is being captured into a
    // graph). Mode 2: a one-thread kernel that stores the device wall clock into the next
    // slot of the caller's stamp buffer, for calls run or captured (every replay of a graph
    // writes them again). Never past kMaxTimingEvents records / the buffer.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (ctx->timing) {
        const hipError_t ce = hipStreamIsCapturing(st, &cap);
        if (ce != hipSuccess) return hip_fail(ctx, ce, "stream capture status");
    }
    const bool capturing = cap != hipStreamCaptureStatusNone;
    const bool timed = (ctx->timing == 1 && !capturing) || ctx->timing == 2;
    if (timed && ctx->used + 4 * (size_t)L.n_chunks > kMaxTimingEvents) ctx->overflow = true;
    if (timed && ctx->timing == 2 && (int64_t)(ctx->used + 4 * (size_t)L.n_chunks) > ctx->stamp_cap)
        ctx->overflow = true;
    auto mark = [&]() -> hipError_t {
        if (!timed || ctx->overflow) return hipSuccess;
        if (ctx->timing == 2) {
            hipLaunchKernelGGL(wall_stamp_kernel, dim3(1), dim3(1), 0, st, ctx->stamps + ctx->used++);
            return hipSuccess;
        }
        if (ctx->used == ctx->pool.size()) {
            hipEvent_t e;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ctx->pool.push_back(e);
        }
        return hipEventRecord(ctx->pool[ctx->used++], st);
    };
    for (int64_t ci = 0; ci < L.n_chunks && b->total_periods > 0; ++ci) {
        p.p0 = ci * L.chunk;
        p.chunk = std::min<int64_t>(L.chunk, b->total_periods - p.p0);
        p.counters = counters + (int64_t)kChunkCounters * ci;
        const int64_t g1 = std::min<int64_t>(L.n_slots, p.chunk);
        hipError_t te = mark();
        hipLaunchKernelGGL(prep_kernel<Real>, dim3((unsigned)g1), dim3(kBlock), 0, st, p);
        if (te == hipSuccess) te = mark();
        const int64_t g2 = std::min<int64_t>(L.n_win, (p.chunk * L.kmax + 3) / 4);
        hipLaunchKernelGGL(windows_kernel<Real>, dim3((unsigned)std::max<int64_t>(1, g2)),
                           dim3(kBlock), 0, st, p);
        if (te == hipSuccess) te = mark();
        hipLaunchKernelGGL(reduce_kernel<Real>, dim3((unsigned)((p.chunk + tpb - 1) / tpb)),
                           dim3(tpb), 0, st, p);
        if (te == hipSuccess) te = mark();
        if (te != hipSuccess) return hip_fail(ctx, te, "timing event");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "search launch");
    return GTLS_OK;
}

int occupancy_slots(gtls_ctx* ctx, const void* kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return per_cu * ctx->n_cu;
}

// Running-median buffer of gtls_spectra_batched_noalloc: one copy per GPU, held until
// the process ends; its host-side name is not a device address (gtls_create looks up the
// device address of each GPU's copy).
constexpr int64_t kNoAllocPeriods = 1LL << 22;   // 32 MiB of float64
static __device__ __align__(256) double scratch[kNoAllocPeriods];

// The spectra kernels of both spectra entry points; med: total_periods doubles.
int launch_spectra(gtls_ctx* ctx, int32_t n_curves, const int64_t* per_off, const double* chi2,
                   int64_t total_periods, int32_t kernel, double* med, double* power_raw,
                   double* power, double* sde_raw, double* sde, int32_t* status) {
    SpecParams p;
    p.n_curves = n_curves;
    p.per_off = per_off;
    p.chi2 = chi2;
    p.kernel = kernel;
    p.power_raw = power_raw;
    p.power = power;
    p.sde_raw = sde_raw;
    p.sde = sde;
    p.status = status;
    hipLaunchKernelGGL(spectra_pass1_kernel, dim3(n_curves), dim3(kSpecBlock), 0, ctx->stream, p);
    if (total_periods > 0)
        hipLaunchKernelGGL(running_median_kernel,
                           dim3((unsigned)((total_periods + kSpecBlock - 1) / kSpecBlock)),
                           dim3(kSpecBlock), 0, ctx->stream, p, total_periods, med);
    hipLaunchKernelGGL(spectra_pass2_kernel, dim3(n_curves), dim3(kSpecBlock), 0, ctx->stream, p,
                       (const double*)med);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "spectra launch");
    return GTLS_OK;
}

int spectra_args_ok(gtls_ctx* ctx, int32_t n_curves, const int64_t* per_off, const double* chi2,
                    int64_t total_periods, int32_t kernel, const double* power_raw,
                    const double* power, const double* sde_raw, const double* sde,
                    const int32_t* status) {
    if (n_curves <= 0 || total_periods < 0 || !per_off || !chi2 || !power_raw || !power ||
        !sde_raw || !sde || !status)
        return fail(ctx, GTLS_E_SHAPE, "bad spectra arguments");
    if (kernel < 1 || kernel > kMaxMedianKernel || kernel % 2 == 0)
        return fail(ctx, GTLS_E_SHAPE, "median kernel must be odd and <= 255");
    return GTLS_OK;
}

}  // namespace

extern "C" {

int gtls_create(int device, void* stream, gtls_ctx** out) {
    if (!out) return GTLS_E_SHAPE;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return -(int)e;
    if (device < 0 || device >= count) return GTLS_E_DEVICE;
    DeviceGuard guard(device);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return -(int)e;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GTLS_E_DEVICE;
    gtls_ctx* ctx = new gtls_ctx();
    ctx->device = device;
    ctx->stream = (hipStream_t)stream;
    ctx->n_cu = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess)
        ctx->wall_khz = (double)khz;
    // this device's copy of the no-allocation spectra buffer (the device is current here)
    e = hipGetSymbolAddress((void**)&ctx->spectra_scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        delete ctx;
        return -(int)e;
    }
    // Each period chunk ends with the tails of its kernels, so chunks are made large: up
    // to 64 GiB of per-period buffers, at most a quarter of the device memory.
    ctx->chunk_bytes = std::min<int64_t>(64LL << 30, (int64_t)(prop.totalGlobalMem / 4));
    if (const char* e = std::getenv("GTLS_MFMA")) ctx->use_mfma = std::atoi(e) != 0;
    if (const char* e = std::getenv("GTLS_TOEP_Q"))
        ctx->toep_q = std::max(0, std::min(kToepQ, std::atoi(e)));
    if (const char* e = std::getenv("GTLS_CHUNK_MB")) {
        const long long mb = std::atoll(e);
        if (mb > 0) ctx->chunk_bytes = mb << 20;
    }
    ctx->prep_fp64 = occupancy_slots(ctx, (const void*)prep_kernel<double>);
    ctx->prep_fp32 = occupancy_slots(ctx, (const void*)prep_kernel<float>);
    ctx->win_fp64 = occupancy_slots(ctx, (const void*)windows_kernel<double>);
    ctx->win_fp32 = occupancy_slots(ctx, (const void*)windows_kernel<float>);
    ctx->slots_sort = occupancy_slots(ctx, (const void*)fold_sort_kernel);
    *out = ctx;
    return GTLS_OK;
}

void gtls_destroy(gtls_ctx* ctx) { delete ctx; }

int gtls_set_stream(gtls_ctx* ctx, void* stream) {
    if (!ctx) return GTLS_E_SHAPE;
    ctx->stream = (hipStream_t)stream;
    return GTLS_OK;
}

const char* gtls_last_error(const gtls_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int gtls_set_tuning(gtls_ctx* ctx, int32_t toep_q, int64_t chunk_mb, int32_t wg_per_cu,
                    int32_t quad_q) {
    if (!ctx) return GTLS_E_SHAPE;
    if (toep_q < 0 || toep_q > kToepQ) return fail(ctx, GTLS_E_SHAPE, "toep_q out of [0, 24]");
    if (chunk_mb < 0 || chunk_mb > (INT64_MAX >> 20))
        return fail(ctx, GTLS_E_SHAPE, "chunk_mb out of [0, 2^43)");
    if (wg_per_cu < 0) return fail(ctx, GTLS_E_SHAPE, "wg_per_cu < 0");
    if (quad_q < 0) return fail(ctx, GTLS_E_SHAPE, "quad_q < 0");
    ctx->wg_per_cu = wg_per_cu;
    ctx->toep_q = toep_q;
    if (chunk_mb > 0) ctx->chunk_bytes = chunk_mb << 20;
    ctx->quad_q = quad_q > 0 ? quad_q : kQuadQ;
    return GTLS_OK;
}

int gtls_set_timing(gtls_ctx* ctx, int32_t on) {
    if (!ctx) return GTLS_E_SHAPE;
    ctx->timing = on != 0;
    ctx->used = 0;
    ctx->overflow = false;
    return GTLS_OK;
}

int64_t gtls_kernel_ms(gtls_ctx* ctx, double* ms) {
    if (!ctx) return GTLS_E_SHAPE;
    if (!ms) return fail(ctx, GTLS_E_SHAPE, "NULL pointer");
    ms[0] = ms[1] = ms[2] = 0.0;
    if (ctx->overflow) {
        ctx->used = 0;
        ctx->overflow = false;
        return fail(ctx, GTLS_E_SHAPE, "more than 16384 period chunks timed since the last "
                                       "gtls_kernel_ms call; records discarded");
    }
    const size_t n = ctx->used / 4 * 4;
    DeviceGuard guard(ctx->device);
    if (ctx->timing == 2) {
        // graph timing: the stamps of the last run or replay
        std::vector<uint64_t> h(n);
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess && n > 0)
            e = hipMemcpy(h.data(), ctx->stamps, n * sizeof(uint64_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(ctx, e, "timing stamps read");
        for (size_t i = 0; i < n; i += 4)
            for (int k = 0; k < 3; ++k)
                ms[k] += (double)(int64_t)(h[i + k + 1] - h[i + k]) / ctx->wall_khz;
        return (int64_t)(n / 4);
    }
    if (n > 0) {
        const hipError_t e = hipEventSynchronize(ctx->pool[n - 1]);
        if (e != hipSuccess) return hip_fail(ctx, e, "timing sync");
    }
    for (size_t i = 0; i < n; i += 4) {
        for (int k = 0; k < 3; ++k) {
            float t = 0.f;
            const hipError_t e = hipEventElapsedTime(&t, ctx->pool[i + k], ctx->pool[i + k + 1]);
            if (e != hipSuccess) return hip_fail(ctx, e, "timing read");
            ms[k] += t;
        }
    }
    ctx->used = 0;
    return (int64_t)(n / 4);
}

int gtls_set_graph_timing(gtls_ctx* ctx, uint64_t* stamps, int64_t max_chunks) {
    if (!ctx) return GTLS_E_SHAPE;
    ctx->used = 0;
    ctx->overflow = false;
    if (!stamps) {
        ctx->timing = 0;
        ctx->stamps = nullptr;
        ctx->stamp_cap = 0;
        return GTLS_OK;
    }
    if (max_chunks <= 0 || max_chunks > (int64_t)kMaxTimingEvents / 4)
        return fail(ctx, GTLS_E_SHAPE, "max_chunks out of [1, 16384]");
    if (!(ctx->wall_khz > 0.0)) return fail(ctx, GTLS_E_DEVICE, "no wall-clock rate");
    ctx->timing = 2;
    ctx->stamps = stamps;
    ctx->stamp_cap = 4 * max_chunks;
    return GTLS_OK;
}

int32_t gtls_search_slots(const gtls_ctx* ctx) { return ctx ? ctx->prep_fp32 : 0; }

int64_t gtls_search_workspace_bytes(gtls_ctx* ctx, int32_t mode, const gtls_batch* b) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C) return fail(ctx, GTLS_E_MODE, "unknown mode");
    std::string why;
    if (!batch_ok(b, why)) return fail(ctx, GTLS_E_SHAPE, why);
    const SearchLayout L = plan_search(ctx, mode, b);
    if (!layout_ok(L, mode, why)) return fail(ctx, GTLS_E_SHAPE, why);
    return L.total;
}

int gtls_search_batched(gtls_ctx* ctx, int32_t mode, const gtls_batch* b, void* workspace,
                        int64_t workspace_bytes, double* chi2, int32_t* best_k,
                        int32_t* best_s) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C) return fail(ctx, GTLS_E_MODE, "unknown mode");
    std::string why;
    if (!batch_ok(b, why)) return fail(ctx, GTLS_E_SHAPE, why);
    if (!workspace || !chi2 || !best_k || !best_s) return fail(ctx, GTLS_E_SHAPE, "NULL output or workspace");
    DeviceGuard guard(ctx->device);
    const SearchLayout L = plan_search(ctx, mode, b);
    if (!layout_ok(L, mode, why)) return fail(ctx, GTLS_E_SHAPE, why);
    if (workspace_bytes < L.total)
        return fail(ctx, GTLS_E_WORKSPACE, "workspace too small: need " + std::to_string(L.total) +
                                               " bytes, got " + std::to_string(workspace_bytes));
    if (mode == GTLS_MODE_FP64)
        return launch_search<double>(ctx, b, (char*)workspace, L, chi2, best_k, best_s);
    return launch_search<float>(ctx, b, (char*)workspace, L, chi2, best_k, best_s);
}

// Single curve: its CSR offsets are written into the head of the workspace by a tiny
// device kernel, so the call stays free of host-to-device copies.
namespace {
__global__ void single_offsets_kernel(int64_t* o, int64_t n, int64_t np, int64_t nw) {
    o[0] = 0; o[1] = n;
    o[2] = 0; o[3] = np;
    o[4] = 0; o[5] = nw;
}
constexpr int64_t kSingleHead = 256;
}  // namespace

int64_t gtls_search_workspace_bytes_1(gtls_ctx* ctx, int32_t mode, int64_t n,
                                      int64_t n_periods, int32_t n_widths,
                                      int64_t total_tmpl, int32_t max_width) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C) return fail(ctx, GTLS_E_MODE, "unknown mode");
    if (n <= 0 || n_periods < 0 || n_widths <= 0 || total_tmpl < 0 || max_width < 0 ||
        n > (1LL << 30))
        return fail(ctx, GTLS_E_SHAPE, "sizes out of range");
    gtls_batch b;
    std::memset(&b, 0, sizeof(b));
    b.n_curves = 1;
    b.total_points = n;
    b.total_periods = n_periods;
    b.total_widths = n_widths;
    b.total_tmpl = total_tmpl;
    b.max_points = (int32_t)n;
    b.max_width = max_width;
    b.max_period_widths = n_widths;
    return kSingleHead + plan_search(ctx, mode, &b).total;
}

int gtls_search(gtls_ctx* ctx, int32_t mode, const double* t, const double* y,
                const double* dy, int64_t n, const double* periods, int64_t n_periods,
                const int32_t* k_lo, const int32_t* k_hi, const int32_t* widths,
                int32_t n_widths, int32_t max_width, const double* overshoot,
                const int64_t* tmpl_off, const double* tmpl, int64_t total_tmpl,
                int64_t skip_point, double depth_min, int32_t negative_time, void* workspace,
                int64_t workspace_bytes, double* chi2, int32_t* best_k, int32_t* best_s) {
    if (!ctx) return GTLS_E_SHAPE;
    int64_t need = gtls_search_workspace_bytes_1(ctx, mode, n, n_periods, n_widths, total_tmpl,
                                                 max_width);
    if (need < 0) return (int)need;
    if (!workspace) return fail(ctx, GTLS_E_SHAPE, "NULL workspace");
    if (workspace_bytes < need)
        return fail(ctx, GTLS_E_WORKSPACE, "workspace too small: need " + std::to_string(need) +
                                               " bytes, got " + std::to_string(workspace_bytes));
    DeviceGuard guard(ctx->device);
    int64_t* head = (int64_t*)workspace;
    hipLaunchKernelGGL(single_offsets_kernel, dim3(1), dim3(1), 0, ctx->stream, head, n,
                       n_periods, (int64_t)n_widths);
    gtls_batch b;
    std::memset(&b, 0, sizeof(b));
    b.n_curves = 1;
    b.pt_off = head;
    b.t = t;
    b.y = y;
    b.dy = dy;
    b.per_off = head + 2;
    b.periods = periods;
    b.k_lo = k_lo;
    b.k_hi = k_hi;
    b.wid_off = head + 4;
    b.widths = widths;
    b.overshoot = overshoot;
    b.tmpl_off = tmpl_off;
    b.tmpl = tmpl;
    b.skip_point = skip_point;
    b.depth_min = depth_min;
    b.negative_time = negative_time;
    b.total_points = n;
    b.total_periods = n_periods;
    b.total_widths = n_widths;
    b.total_tmpl = total_tmpl;
    b.max_points = (int32_t)n;
    b.max_width = max_width;
    b.max_period_widths = n_widths;
    return gtls_search_batched(ctx, mode, &b, (char*)workspace + kSingleHead,
                               workspace_bytes - kSingleHead, chi2, best_k, best_s);
}

int64_t gtls_fold_sort_workspace_bytes(gtls_ctx* ctx, int64_t n) {
    if (!ctx) return GTLS_E_SHAPE;
    if (n <= 0 || n > (1LL << 30)) return fail(ctx, GTLS_E_SHAPE, "n out of range");
    return fold_sort_slot_bytes(n) * ctx->slots_sort;
}

int gtls_fold_sort(gtls_ctx* ctx, const double* t, int64_t n, const double* periods,
                   int64_t n_periods, int32_t negative_time, void* workspace,
                   int64_t workspace_bytes, int32_t* perm) {
    if (!ctx) return GTLS_E_SHAPE;
    if (!t || !periods || !perm || !workspace) return fail(ctx, GTLS_E_SHAPE, "NULL pointer");
    int64_t need = gtls_fold_sort_workspace_bytes(ctx, n);
    if (need < 0) return (int)need;
    if (n_periods < 0) return fail(ctx, GTLS_E_SHAPE, "n_periods < 0");
    if (workspace_bytes < need) return fail(ctx, GTLS_E_WORKSPACE, "workspace too small");
    if (n_periods == 0) return GTLS_OK;
    DeviceGuard guard(ctx->device);
    const int64_t grid = std::min<int64_t>(ctx->slots_sort, n_periods);
    hipLaunchKernelGGL(fold_sort_kernel, dim3((unsigned)grid), dim3(kBlock), 0, ctx->stream, t,
                       n, periods, n_periods, negative_time, (char*)workspace, perm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "fold_sort launch");
    return GTLS_OK;
}

int64_t gtls_spectra_workspace_bytes(gtls_ctx* ctx, int64_t total_periods) {
    if (!ctx) return GTLS_E_SHAPE;
    if (total_periods < 0) return fail(ctx, GTLS_E_SHAPE, "total_periods < 0");
    return align_up(8 * (total_periods > 0 ? total_periods : 1));
}

int gtls_spectra_batched(gtls_ctx* ctx, int32_t n_curves, const int64_t* per_off,
                         const double* chi2, int64_t total_periods, int32_t kernel,
                         void* workspace, int64_t workspace_bytes, double* power_raw,
                         double* power, double* sde_raw, double* sde, int32_t* status) {
    if (!ctx) return GTLS_E_SHAPE;
    const int ok = spectra_args_ok(ctx, n_curves, per_off, chi2, total_periods, kernel,
                                   power_raw, power, sde_raw, sde, status);
    if (ok != GTLS_OK) return ok;
    if (!workspace) return fail(ctx, GTLS_E_SHAPE, "NULL workspace");
    if (workspace_bytes < gtls_spectra_workspace_bytes(ctx, total_periods))
        return fail(ctx, GTLS_E_WORKSPACE, "workspace too small");
    DeviceGuard guard(ctx->device);
    return launch_spectra(ctx, n_curves, per_off, chi2, total_periods, kernel,
                          (double*)workspace, power_raw, power, sde_raw, sde, status);
}

int gtls_spectra_batched_noalloc(gtls_ctx* ctx, int32_t n_curves, const int64_t* per_off,
                                 const double* chi2, int64_t total_periods, int32_t kernel,
                                 double* power_raw, double* power, double* sde_raw,
                                 double* sde, int32_t* status) {
    if (!ctx) return GTLS_E_SHAPE;
    const int ok = spectra_args_ok(ctx, n_curves, per_off, chi2, total_periods, kernel,
                                   power_raw, power, sde_raw, sde, status);
    if (ok != GTLS_OK) return ok;
    if (total_periods > kNoAllocPeriods)
        return fail(ctx, GTLS_E_WORKSPACE,
                    "gtls_spectra_batched_noalloc takes at most 4194304 periods per call, got " +
                        std::to_string(total_periods) + " (use gtls_spectra_batched)");
    if (!ctx->spectra_scratch) return fail(ctx, GTLS_E_DEVICE, "no scratch address for the device");
    DeviceGuard guard(ctx->device);
    // every running-median entry the second pass reads is written by this call first, so
    // the buffer needs no clearing
    return launch_spectra(ctx, n_curves, per_off, chi2, total_periods, kernel,
                          ctx->spectra_scratch, power_raw, power, sde_raw, sde, status);
}

int gtls_transit_fwd(gtls_ctx* ctx, int32_t mode, int64_t n_models, const double* theta,
                     const int32_t* model_curve, const int64_t* pt_off, const double* t,
                     const int64_t* out_off, int64_t max_points, void* flux) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C)
        return fail(ctx, GTLS_E_MODE, "unknown mode");
    if (n_models < 0 || n_models > 65535 || max_points < 0)
        return fail(ctx, GTLS_E_SHAPE, "transit model count must be in [0, 65535]");
    if (n_models == 0 || max_points == 0) return GTLS_OK;
    if (!theta || !model_curve || !pt_off || !t || !out_off || !flux)
        return fail(ctx, GTLS_E_SHAPE, "null transit model argument");
    DeviceGuard guard(ctx->device);
    const unsigned bx = (unsigned)std::min<int64_t>((max_points + 1023) / 1024, 4096);
    const dim3 grid(bx, (unsigned)n_models);
    if (mode == GTLS_MODE_FP64)
        hipLaunchKernelGGL(transit_fwd_kernel<double>, grid, dim3(256), 0, ctx->stream, theta,
                           model_curve, pt_off, t, out_off, (double*)flux);
    else
        hipLaunchKernelGGL(transit_fwd_kernel<float>, grid, dim3(256), 0, ctx->stream, theta,
                           model_curve, pt_off, t, out_off, (float*)flux);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "transit forward launch");
    return GTLS_OK;
}

int gtls_transit_bwd(gtls_ctx* ctx, int32_t mode, int64_t n_models, const double* theta,
                     const int32_t* model_curve, const int64_t* pt_off, const double* t,
                     const int64_t* out_off, const void* grad_flux, double* grad_theta) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C)
        return fail(ctx, GTLS_E_MODE, "unknown mode");
    if (n_models < 0 || n_models > (1LL << 31) - 1)
        return fail(ctx, GTLS_E_SHAPE, "bad transit model count");
    if (n_models == 0) return GTLS_OK;
    if (!theta || !model_curve || !pt_off || !t || !out_off || !grad_flux || !grad_theta)
        return fail(ctx, GTLS_E_SHAPE, "null transit model argument");
    DeviceGuard guard(ctx->device);
    if (mode == GTLS_MODE_FP64)
        hipLaunchKernelGGL(transit_bwd_kernel<double>, dim3((unsigned)n_models), dim3(256), 0,
                           ctx->stream, theta, model_curve, pt_off, t, out_off,
                           (const double*)grad_flux, grad_theta);
    else
        hipLaunchKernelGGL(transit_bwd_kernel<float>, dim3((unsigned)n_models), dim3(256), 0,
                           ctx->stream, theta, model_curve, pt_off, t, out_off,
                           (const float*)grad_flux, grad_theta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "transit backward launch");
    return GTLS_OK;
}

int gtls_transit_chi2(gtls_ctx* ctx, int32_t mode, int64_t n_models, const double* theta,
                      const int32_t* model_curve, const int64_t* pt_off, const double* t,
                      const double* y, const double* w, double* chi2, double* grad_theta) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C)
        return fail(ctx, GTLS_E_MODE, "unknown mode");
    if (n_models < 0 || n_models > (1LL << 31) - 1)
        return fail(ctx, GTLS_E_SHAPE, "bad transit model count");
    if (n_models == 0) return GTLS_OK;
    if (!theta || !model_curve || !pt_off || !t || !y || !w || !chi2)
        return fail(ctx, GTLS_E_SHAPE, "null transit model argument");
    DeviceGuard guard(ctx->device);
    const dim3 grid((unsigned)n_models);
    if (mode == GTLS_MODE_FP64) {
        const dim3 block(chi2_block<double>());
        if (grad_theta)
            hipLaunchKernelGGL((transit_chi2_kernel<double, true>), grid, block, 0, ctx->stream,
                               theta, model_curve, pt_off, t, y, w, chi2, grad_theta);
        else
            hipLaunchKernelGGL((transit_chi2_kernel<double, false>), grid, block, 0, ctx->stream,
                               theta, model_curve, pt_off, t, y, w, chi2, grad_theta);
    } else {
        const dim3 block(chi2_block<float>());
        if (grad_theta)
            hipLaunchKernelGGL((transit_chi2_kernel<float, true>), grid, block, 0, ctx->stream,
                               theta, model_curve, pt_off, t, y, w, chi2, grad_theta);
        else
            hipLaunchKernelGGL((transit_chi2_kernel<float, false>), grid, block, 0, ctx->stream,
                               theta, model_curve, pt_off, t, y, w, chi2, grad_theta);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "transit chi2 launch");
    return GTLS_OK;
}

int64_t gtls_transit_tile_points(void) { return kChi2Tile; }

int gtls_transit_chi2_grouped(gtls_ctx* ctx, int32_t mode, int32_t n_curves, int64_t n_models,
                              int32_t max_models, const double* theta, const int64_t* model_off,
                              const int64_t* pt_off, const int64_t* tile_off, int64_t n_tiles,
                              const double* t, const double* y, const double* w, double* part,
                              int64_t part_len, double* chi2, double* grad_theta) {
    if (!ctx) return GTLS_E_SHAPE;
    if (mode != GTLS_MODE_FP64 && mode != GTLS_MODE_FP32C)
        return fail(ctx, GTLS_E_MODE, "unknown mode");
    if (n_curves <= 0 || n_models < 0 || n_models > (1LL << 31) - 1 || n_tiles < 0 ||
        n_tiles > (1LL << 31) - 1 || max_models < 0)
        return fail(ctx, GTLS_E_SHAPE, "bad curve, model or tile count");
    if (n_models == 0) return GTLS_OK;
    if (max_models < 1) return fail(ctx, GTLS_E_SHAPE, "max_models must be positive");
    if (!theta || !model_off || !pt_off || !tile_off || !t || !y || !w || !chi2)
        return fail(ctx, GTLS_E_SHAPE, "null transit model argument");
    const int sums = grad_theta ? 7 : 1;
    if (n_tiles > 0 && (!part || part_len < n_tiles * ((int64_t)max_models * sums + 1)))
        return fail(ctx, GTLS_E_WORKSPACE,
                    "tile sums need n_tiles * (max_models * (7 or 1) + 1) doubles");
    double* flat = part + n_tiles * (int64_t)max_models * sums;
    DeviceGuard guard(ctx->device);
    const dim3 tiles((unsigned)n_tiles), block(kChi2Threads);
    if (n_tiles > 0) {
        if (mode == GTLS_MODE_FP64) {
            if (grad_theta)
                hipLaunchKernelGGL((transit_chi2_tile_kernel<double, true>), tiles, block, 0,
                                   ctx->stream, theta, model_off, pt_off, tile_off, n_curves,
                                   max_models, t, y, w, part, flat);
            else
                hipLaunchKernelGGL((transit_chi2_tile_kernel<double, false>), tiles, block, 0,
                                   ctx->stream, theta, model_off, pt_off, tile_off, n_curves,
                                   max_models, t, y, w, part, flat);
        } else {
            if (grad_theta)
                hipLaunchKernelGGL((transit_chi2_tile_kernel<float, true>), tiles, block, 0,
                                   ctx->stream, theta, model_off, pt_off, tile_off, n_curves,
                                   max_models, t, y, w, part, flat);
            else
                hipLaunchKernelGGL((transit_chi2_tile_kernel<float, false>), tiles, block, 0,
                                   ctx->stream, theta, model_off, pt_off, tile_off, n_curves,
                                   max_models, t, y, w, part, flat);
        }
    }
    const int64_t nsum = n_models * sums;
    const dim3 sgrid((unsigned)((nsum + 255) / 256)), sblock(256);
    if (grad_theta)
        hipLaunchKernelGGL(transit_chi2_sum_kernel<true>, sgrid, sblock, 0, ctx->stream,
                           model_off, tile_off, n_curves, max_models, n_models, part, flat,
                           chi2, grad_theta);
    else
        hipLaunchKernelGGL(transit_chi2_sum_kernel<false>, sgrid, sblock, 0, ctx->stream,
                           model_off, tile_off, n_curves, max_models, n_models, part, flat,
                           chi2, grad_theta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "grouped transit chi2 launch");
    return GTLS_OK;
}

}  // extern "C"
