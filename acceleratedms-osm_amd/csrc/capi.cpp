// C-ABI implementation: design wrappers, the device plan (MSOSM_GPU_BATCH device pipeline) and the
// fold object (Fold_GPU). Host-side only; device code lives in kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "kernels.h"
#include "msosm_amd.h"
#include "plan_internal.h"

namespace msosm {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

} // namespace msosm

using namespace msosm;

#define HIPCHK(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return set_error(MSOSM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " (" + __FILE__ +      \
                                                ":" + std::to_string(__LINE__) + ")");                                 \
    } while (0)

static int ilog2(long long v)
{
    int r = -1;
    while (v) {
        v >>= 1;
        ++r;
    }
    return r;
}

static bool is_pow2(long long v) { return v > 0 && (v & (v - 1)) == 0; }

extern "C" const char *msosm_last_error(void) { return g_last_error.c_str(); }

extern "C" int msosm_version(void) { return 100; }

// ------------------------------------------------------------------------------------------------
// design
// ------------------------------------------------------------------------------------------------
extern "C" int msosm_design_nd(float bw, float f0, float dm, int *nd)
{
    if (!nd || !(bw > 0) || !(f0 > 0) || !(dm >= 0)) return set_error(MSOSM_ERR_ARG, "design_nd: need bw>0, f0>0, dm>=0");
    *nd = design_nd(bw, f0, dm);
    return MSOSM_OK;
}

extern "C" int msosm_design_min_order(int nd, int *M)
{
    if (!M || nd < 0) return set_error(MSOSM_ERR_ARG, "design_min_order: nd must be >= 0");
    *M = nd == 0 ? 0 : design_min_order(nd);
    return MSOSM_OK;
}

extern "C" int msosm_design_osm_order(unsigned long nd, unsigned long *M)
{
    if (!M) return set_error(MSOSM_ERR_ARG, "design_osm_order: null output");
    *M = design_osm_order(nd);
    return MSOSM_OK;
}

extern "C" int msosm_design_delaycount(int nd, int M, int *dc)
{
    if (!dc || M <= 0 || nd < 0) return set_error(MSOSM_ERR_ARG, "design_delaycount: need M>0, nd>=0");
    *dc = design_delaycount(nd, M);
    return MSOSM_OK;
}

extern "C" int msosm_design_segmentation(float bw, float f0, float dm, int M, int *delay_points, int *delaycount)
{
    if (!delay_points || !is_pow2(M) || !(bw > 0) || !(f0 > 0) || !(dm >= 0))
        return set_error(MSOSM_ERR_ARG, "design_segmentation: need M a power of two, bw>0, f0>0, dm>=0");
    const int dc = design_segmentation(bw, f0, dm, M, delay_points);
    for (long k = 0; k < 2L * M; ++k)
        if (delay_points[k] < 0)
            return set_error(MSOSM_ERR_ARG, "design_segmentation: bins left without a delay (inconsistent design)");
    if (delaycount) *delaycount = dc;
    return MSOSM_OK;
}

extern "C" int msosm_design_build_segments(const int *delay, int fft_len, int max_len, int *seg, int cap, int *count)
{
    if (!delay || !seg || !count || fft_len < 0 || max_len <= 0)
        return set_error(MSOSM_ERR_ARG, "design_build_segments: bad arguments");
    const int n = design_build_segments(delay, fft_len, max_len, seg, cap);
    if (n < 0) return set_error(MSOSM_ERR_ARG, "design_build_segments: cap too small");
    *count = n;
    return MSOSM_OK;
}

extern "C" int msosm_design_transfer_function(float bw, float f0, float dm, int M, int phase_mode, float *H)
{
    if (!H || !is_pow2(M) || M < 2 || (phase_mode != MSOSM_PHASE_FP64 && phase_mode != MSOSM_PHASE_FP32))
        return set_error(MSOSM_ERR_ARG, "design_transfer_function: need M a power of two >= 2 and a valid phase mode");
    return design_transfer_function(bw, f0, dm, M, phase_mode, H, nullptr);
}

extern "C" int msosm_design_transfer_functions(float bw, float f0, const float *dms, int num_dms, int M,
                                               int phase_mode, float *H, int nthreads)
{
    if (!H || !dms || num_dms <= 0 || !is_pow2(M) || M < 2 ||
        (phase_mode != MSOSM_PHASE_FP64 && phase_mode != MSOSM_PHASE_FP32))
        return set_error(MSOSM_ERR_ARG, "design_transfer_functions: bad arguments");
    return design_transfer_functions(bw, f0, dms, num_dms, M, phase_mode, H, nthreads);
}

extern "C" int msosm_design_transfer_function_ddm(float bw, float f0, float dm, int M, int phase_mode, float *dH)
{
    if (!dH || !is_pow2(M) || M < 2 || (phase_mode != MSOSM_PHASE_FP64 && phase_mode != MSOSM_PHASE_FP32))
        return set_error(MSOSM_ERR_ARG,
                         "design_transfer_function_ddm: need M a power of two >= 2 and a valid phase mode");
    return design_transfer_function(bw, f0, dm, M, phase_mode, dH, nullptr, 1);
}

// ------------------------------------------------------------------------------------------------
// fold object
// ------------------------------------------------------------------------------------------------
struct msosm_fold {
    int device = 0;
    double period = 0, fs = 0, dt = 0, t_start = 0;
    unsigned long size = 0, nbins = 0;
    int S = 1;
    double *sums = nullptr;
    unsigned long long *cnts = nullptr;
    float *prof = nullptr;
};

static int fold_free(msosm_fold *f)
{
    if (!f) return MSOSM_OK;
    (void)hipSetDevice(f->device);
    if (f->sums) (void)hipFree(f->sums);
    if (f->cnts) (void)hipFree(f->cnts);
    if (f->prof) (void)hipFree(f->prof);
    delete f;
    return MSOSM_OK;
}

extern "C" int msosm_fold_create(double period, double fs, unsigned long size, unsigned long nbins, int num_series,
                                 int device, msosm_fold **out)
{
    if (!out || !(period > 0) || !(fs > 0) || size == 0 || num_series <= 0)
        return set_error(MSOSM_ERR_ARG, "fold_create: need period>0, fs>0, size>0, num_series>0");
    if (nbins == 0) {
        // time_bin = -1 -> period_samples, computed as (unsigned long)(float period * float fs) (fold_gpu.cpp:10-15)
        const float ps = (float)period * (float)fs;
        nbins = (unsigned long)ps;
        if (nbins == 0) return set_error(MSOSM_ERR_ARG, "fold_create: period*fs < 1 sample");
    }
    msosm_fold *f = new msosm_fold();
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) device = 0;
    }
    f->device = device;
    f->period = period;
    f->fs = fs;
    f->dt = 1.0 / fs; // fold_gpu.cpp:50
    f->size = size;
    f->nbins = nbins;
    f->S = num_series;
    const size_t n = nbins * (size_t)num_series;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(&f->sums, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&f->cnts, n * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&f->prof, n * sizeof(float));
    if (e == hipSuccess) e = hipMemset(f->sums, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(f->cnts, 0, n * sizeof(unsigned long long));
    if (e != hipSuccess) {
        fold_free(f);
        return set_error(MSOSM_ERR_ALLOC, std::string("fold_create: ") + hipGetErrorString(e));
    }
    *out = f;
    return MSOSM_OK;
}

extern "C" int msosm_fold_destroy(msosm_fold *f) { return fold_free(f); }

extern "C" int msosm_fold_get_nbins(msosm_fold *f, unsigned long *nbins)
{
    if (!f || !nbins) return set_error(MSOSM_ERR_ARG, "fold_get_nbins: null argument");
    *nbins = f->nbins;
    return MSOSM_OK;
}

extern "C" int msosm_fold_accumulate(msosm_fold *f, const void *pol1, const void *pol2, long long stride, void *stream)
{
    if (!f || !pol1) return set_error(MSOSM_ERR_ARG, "fold_accumulate: null fold or series");
    if (f->S > 1 && stride < (long long)f->size)
        return set_error(MSOSM_ERR_ARG, "fold_accumulate: series stride shorter than the fold size");
    HIPCHK(hipSetDevice(f->device));
    HIPCHK(launch_fold_series(static_cast<const float2 *>(pol1), static_cast<const float2 *>(pol2),
                              (long long)f->size, stride, f->S, f->nbins, f->t_start, f->dt, f->period, f->sums,
                              f->cnts, static_cast<hipStream_t>(stream)));
    f->t_start += static_cast<double>(f->size) * f->dt; // fold_gpu.cpp:53
    return MSOSM_OK;
}

extern "C" int msosm_stream_wait(void *waiting, void *producer)
{
    if (waiting == producer) return MSOSM_OK;
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev, static_cast<hipStream_t>(producer));
    if (e == hipSuccess) e = hipStreamWaitEvent(static_cast<hipStream_t>(waiting), ev, 0);
    if (ev) (void)hipEventDestroy(ev); // the wait keeps what it needs
    if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("stream_wait: ") + hipGetErrorString(e));
    return MSOSM_OK;
}

extern "C" int msosm_fold_read(msosm_fold *f, float *profile, double *sums, unsigned long long *counts)
{
    if (!f) return set_error(MSOSM_ERR_ARG, "fold_read: null fold");
    HIPCHK(hipSetDevice(f->device));
    const size_t n = f->nbins * (size_t)f->S;
    HIPCHK(hipDeviceSynchronize());
    if (profile) {
        HIPCHK(launch_profile(f->sums, f->cnts, f->prof, (long long)n, nullptr));
        HIPCHK(hipMemcpy(profile, f->prof, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (sums) HIPCHK(hipMemcpy(sums, f->sums, n * sizeof(double), hipMemcpyDeviceToHost));
    if (counts) HIPCHK(hipMemcpy(counts, f->cnts, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return MSOSM_OK;
}

extern "C" int msosm_fold_reset(msosm_fold *f)
{
    if (!f) return set_error(MSOSM_ERR_ARG, "fold_reset: null fold");
    HIPCHK(hipSetDevice(f->device));
    HIPCHK(hipDeviceSynchronize());
    const size_t n = f->nbins * (size_t)f->S;
    HIPCHK(hipMemset(f->sums, 0, n * sizeof(double)));
    HIPCHK(hipMemset(f->cnts, 0, n * sizeof(unsigned long long)));
    f->t_start = 0.0;
    return MSOSM_OK;
}

// ------------------------------------------------------------------------------------------------
// fold without allocation: the partial sums and counts in a static device array
// ------------------------------------------------------------------------------------------------
// [0, MAX) FP64 sums, [MAX, 2 MAX) u64 counts. Each GPU holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns on that device.
static __device__ __align__(256) unsigned long long scratch[2 * MSOSM_FOLD_NOALLOC_MAX_ENTRIES];

static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr; // per device, from hipGetSymbolAddress with that device current

// the device address of `scratch` on the current device (looked up once per device)
static int scratch_address(int device, unsigned long long **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(MSOSM_ERR_HIP, std::string("fold_series_noalloc: hipGetSymbolAddress: ") +
                                                hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<unsigned long long *>(g_scratch_ptr[device]);
    return MSOSM_OK;
}

extern "C" int msosm_fold_series_noalloc(const void *pol1, const void *pol2, long long size, long long stride,
                                         int num_series, unsigned long nbins, double t_start, double fs,
                                         double period, float *profile, unsigned long long *counts, void *stream)
{
    if (!pol1 || !profile || size <= 0 || num_series <= 0 || nbins == 0 || !(fs > 0) || !(period > 0))
        return set_error(MSOSM_ERR_ARG, "fold_series_noalloc: need series, a profile, size > 0, num_series > 0, "
                                        "nbins > 0, fs > 0 and period > 0");
    if (num_series > 1 && stride < size)
        return set_error(MSOSM_ERR_ARG, "fold_series_noalloc: series stride shorter than the series size");
    const unsigned long long n = (unsigned long long)nbins * (unsigned long long)num_series;
    if (n > MSOSM_FOLD_NOALLOC_MAX_ENTRIES)
        return set_error(MSOSM_ERR_ARG, "fold_series_noalloc: num_series * nbins = " + std::to_string(n) +
                                            " exceeds the static scratch of " +
                                            std::to_string(MSOSM_FOLD_NOALLOC_MAX_ENTRIES) + " bins");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    unsigned long long *ptr = nullptr;
    if (scratch_address(dev, &ptr) != MSOSM_OK) return MSOSM_ERR_HIP;
    double *sums = reinterpret_cast<double *>(ptr);
    unsigned long long *cnts = ptr + MSOSM_FOLD_NOALLOC_MAX_ENTRIES;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the previous call's partials are cleared by a fill kernel (no memset node under graph capture)
    HIPCHK(launch_fill(sums, (long long)(n * sizeof(double)), 0u, st));
    HIPCHK(launch_fill(cnts, (long long)(n * sizeof(unsigned long long)), 0u, st));
    // Fold_GPU::fold_data_phase (fold_gpu.cpp:48-54) with dt = 1/fs, then computeFoldedData (:56-60)
    HIPCHK(launch_fold_series(static_cast<const float2 *>(pol1), static_cast<const float2 *>(pol2), size, stride,
                              num_series, nbins, t_start, 1.0 / fs, period, sums, cnts, st));
    HIPCHK(launch_profile(sums, cnts, profile, (long long)n, st, counts));
    return MSOSM_OK;
}

// ------------------------------------------------------------------------------------------------
// plan
// ------------------------------------------------------------------------------------------------
struct EventPair {
    hipEvent_t a, b;
};

struct msosm_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int M = 0, L = 0, log2L = 0, D = 0, B = 0, dc = 0, FBS = 0, in_fmt = 0, out_flags = 0, G = 0;
    size_t in_bytes = 0;
    void *in_dev[2] = {nullptr, nullptr}; // staging buffers for host inputs
    hipEvent_t in_ev[2] = {nullptr, nullptr}; // staging slot read by its push's forward transform
    // copies beside the compute stream (msosm_gpu_batch.cpp:15-27,203-245): host input is staged on
    // in_stream, outputs leave on out_stream; the uint16 output has two buffers (push parity), so one
    // push's device-to-host copy runs under the next push's kernels
    hipStream_t in_stream = nullptr, out_stream = nullptr;
    hipEvent_t in_done[2] = {nullptr, nullptr};    // staging slot filled
    hipEvent_t out_ready[2] = {nullptr, nullptr};  // the push of that parity has written its outputs
    hipEvent_t d2h_u16[2] = {nullptr, nullptr};    // uint16 buffer of that parity copied out
    hipEvent_t d2h_c64 = nullptr, d2h_f16 = nullptr; // the single complex64 / FP16 buffers copied out
    uint32_t *out_u16b[2] = {nullptr, nullptr};
    long long pushes = 0;
    float2 *hist = nullptr, *ring = nullptr, *H = nullptr, *scratch = nullptr, *out_c64 = nullptr, *tw = nullptr;
    uint16_t *delay = nullptr, *bins = nullptr;
    uint32_t *out_u16 = nullptr, *out_f16 = nullptr;
    float *out_scale = nullptr;
    int tw_n1 = 0, tw_n2 = 0, tw_lo = 0, tw_hi = 0, tw_s = 0;
    long long g0 = 0;
    int max_fold_bins = 0;
    msosm_fold *fold = nullptr;
    size_t device_bytes = 0;
    bool profile = false;
    std::vector<EventPair> pending[2], pool;
    double prof_ms[2] = {0, 0};
    long long prof_n[2] = {0, 0};
    std::vector<char> dm_set;
    int launch_mode = MSOSM_LAUNCH_PERSISTENT;
    // accumulate launch (lengths 2^14..2^16, delays < 256): the ring also in its layout (perm_pos; the
    // same array when R = 1), H and dH in its chirp layout (acc_chirp_pos) and the delays as bytes in
    // thread order
    bool accum_ok = false;
    int perm_rl = 0, perm_sl = 0;
    // lengths 2^17, 2^18: the accumulate launch runs acc_q sub-units of 2^16 points per unit on the
    // sub-unit spectra a gathering pass writes to xpre, with the sub-unit length's twiddle tables at
    // these offsets of tw (acc_q = 1: one unit, no gathering pass)
    int acc_q = 1, tw_sub_lo = 0, tw_sub_hi = 0, tw_sub_128 = 0, tw_sub_s = 0;
    float2 *xpre = nullptr;
    unsigned int *sub_sync = nullptr; // [D][B] arrival counters of the sub-unit launch
    float2 *ring_p = nullptr, *H_p = nullptr, *Hdot_p = nullptr;
    uint8_t *delay_acc = nullptr;
    int sched_wg = 0, order = MSOSM_ORDER_XCD_DM_MAJOR; // tile schedule (0 workgroups = G)
    // DM-derivative pass: dH/dDM per DM and the per-push primal series it is paired with
    float2 *Hdot = nullptr, *tan_y = nullptr;
    std::vector<char> tan_set;
};

template <class T>
static hipError_t dalloc(msosm_plan *p, T **ptr, size_t bytes)
{
    hipError_t e = hipMalloc((void **)ptr, bytes);
    if (e == hipSuccess) p->device_bytes += bytes;
    return e;
}

static void plan_free(msosm_plan *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->in_stream) (void)hipStreamSynchronize(p->in_stream);
    if (p->out_stream) (void)hipStreamSynchronize(p->out_stream);
    void *bufs[] = {p->in_dev[0], p->in_dev[1], p->hist, p->ring, p->H, p->scratch, p->out_c64, p->tw, p->delay,
                    p->bins, p->out_u16b[0], p->out_u16b[1], p->out_f16, p->out_scale, p->Hdot, p->tan_y, p->ring_p,
                    p->H_p, p->Hdot_p, p->delay_acc, p->xpre, p->sub_sync};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : {p->in_ev[0], p->in_ev[1], p->in_done[0], p->in_done[1], p->out_ready[0], p->out_ready[1],
                         p->d2h_u16[0], p->d2h_u16[1], p->d2h_c64, p->d2h_f16})
        if (e) (void)hipEventDestroy(e);
    if (p->in_stream) (void)hipStreamDestroy(p->in_stream);
    if (p->out_stream) (void)hipStreamDestroy(p->out_stream);
    for (int k = 0; k < 2; ++k)
        for (auto &e : p->pending[k]) p->pool.push_back(e);
    for (auto &e : p->pool) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" int msosm_plan_create(const msosm_plan_desc *d, msosm_plan **out)
{
    if (!d || !out) return set_error(MSOSM_ERR_ARG, "plan_create: null argument");
    *out = nullptr;
    if (!is_pow2(d->M)) return set_error(MSOSM_ERR_ARG, "plan_create: M must be a power of two");
    const int log2L = ilog2(2LL * d->M);
    if (!length_supported(log2L))
        return set_error(MSOSM_ERR_UNSUPPORTED, "plan_create: FFT length 2M=" + std::to_string(2LL * d->M) +
                                                     " outside the supported range [2^12, 2^22]");
    if (d->num_dms <= 0 || d->batch <= 0) return set_error(MSOSM_ERR_ARG, "plan_create: num_dms and batch must be > 0");
    // unit (and sub-unit) positions of the dedispersion launches are 32-bit
    if ((long long)d->num_dms * d->batch * 4 >= (1ll << 31))
        return set_error(MSOSM_ERR_ARG, "plan_create: num_dms * batch must be below 2^29");
    if (d->delaycount <= 0 || d->delaycount > 65535)
        return set_error(MSOSM_ERR_ARG, "plan_create: delaycount must be in [1, 65535]");
    if (d->in_format < MSOSM_IN_U16 || d->in_format > MSOSM_IN_C64)
        return set_error(MSOSM_ERR_ARG, "plan_create: unknown input format");
    if (d->out_flags & ~(MSOSM_OUT_C64 | MSOSM_OUT_U16 | MSOSM_OUT_F16))
        return set_error(MSOSM_ERR_ARG, "plan_create: unknown output flags");

    msosm_plan *p = new msosm_plan();
    int dev = d->device;
    if (dev < 0) {
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    }
    p->device = dev;
    auto fail = [&](int code, const std::string &m) {
        plan_free(p);
        return set_error(code, m);
    };
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return fail(MSOSM_ERR_HIP, std::string("plan_create: hipSetDevice: ") + hipGetErrorString(e));
    p->M = d->M;
    p->L = 2 * d->M;
    p->log2L = log2L;
    p->D = d->num_dms;
    p->B = d->batch;
    p->dc = d->delaycount;
    p->in_fmt = d->in_format;
    p->out_flags = d->out_flags;
    // ring depth: roundup(delaycount + batch - 1, batch) (msosm_gpu_batch.cpp:113-117)
    long long fbs = (long long)p->dc + p->B - 1;
    if (fbs % p->B) fbs += p->B - fbs % p->B;
    if (fbs > 0x7fffffffLL) return fail(MSOSM_ERR_ARG, "plan_create: ring too deep");
    p->FBS = (int)fbs;
    p->in_bytes = p->in_fmt == MSOSM_IN_U16 ? 4 : p->in_fmt == MSOSM_IN_I8 ? 2 : 8;

    int cus = 0, max_lds = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return fail(MSOSM_ERR_HIP, std::string("plan_create: device query: ") + hipGetErrorString(e));
    // kernels use tiles + twiddle tables; the fused fold's histogram (12 B per bin) takes the rest
    const int tiles_lds = tiles_lds_bytes(log2L);
    if (tiles_lds + INV_EXTRA_LDS > max_lds)
        return fail(MSOSM_ERR_UNSUPPORTED, "plan_create: kernels need " + std::to_string(tiles_lds) +
                                               " B of LDS per workgroup, device allows " + std::to_string(max_lds));
    // the accumulate form stores the delays as bytes: blocks 0..255
    // (the sub-unit form of 2^17 and 2^18 has no FP16-pair epilogue: its per-block scale spans sub-units)
    const int acc_q = accum_subunits(log2L);
    p->accum_ok = accum_supported(log2L) && p->dc <= 256 && (unsigned long long)p->FBS * p->L < (1ull << 32) &&
                  (acc_q == 1 || !(p->out_flags & MSOSM_OUT_F16));
    const int acc_lds = p->accum_ok ? accum_lds_bytes(log2L) : 0;
    if (p->accum_ok) {
        p->acc_q = acc_q;
        // sub-unit plans read the natural ring, delays and chirps (in the gathering pass)
        p->perm_rl = acc_q > 1 ? 0 : log2L - 14;
        p->perm_sl = 14;
    }
    p->max_fold_bins = std::min(2048, (max_lds - std::max(tiles_lds, acc_lds) - INV_EXTRA_LDS) / 12);
    e = configure_lds(log2L, tiles_lds, tiles_lds + p->max_fold_bins * 12 + INV_EXTRA_LDS);
    if (e == hipSuccess && acc_lds)
        e = configure_accum_lds(log2L, acc_lds + p->max_fold_bins * 12 + INV_EXTRA_LDS);
    if (e != hipSuccess) return fail(MSOSM_ERR_HIP, std::string("plan_create: LDS attribute: ") + hipGetErrorString(e));
    p->G = std::max(8, (cus / 8) * 8);

    if (d->stream) {
        p->stream = static_cast<hipStream_t>(d->stream);
    } else {
        e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return fail(MSOSM_ERR_HIP, std::string("plan_create: stream: ") + hipGetErrorString(e));
        p->own_stream = true;
    }

    const size_t L = p->L, M = p->M, B = p->B, D = p->D;
    // twiddle tables (FP64 -> complex64): n1, n2, lo, hi
    int n1, n2;
    split_length(log2L, &n1, &n2);
    p->tw_s = (log2L + 1) / 2;
    const int nlo = 1 << p->tw_s, nhi = (int)(L >> p->tw_s);
    // the dedispersion kernel keeps lo and hi in LDS with the sizes of n1 and n2
    if (nlo != n1 || nhi != n2) return fail(MSOSM_ERR_ARG, "plan_create: twiddle split does not match the length split");
    p->tw_n1 = 0;
    p->tw_n2 = n1;
    p->tw_lo = n1 + n2;
    p->tw_hi = n1 + n2 + nlo;
    int ntw = n1 + n2 + nlo + nhi;
    if (p->acc_q > 1) {
        // the sub-unit length 2^16: lo and hi (256 each) split at 8 bits, and the 128-point roots
        p->tw_sub_s = 8;
        p->tw_sub_lo = ntw;
        p->tw_sub_hi = ntw + 256;
        p->tw_sub_128 = ntw + 512;
        ntw += 256 + 256 + 128;
    }
    std::vector<float2> tw(ntw);
    auto put = [&](int off, long k, long n) {
        double re, im;
        unit_root(k, n, &re, &im);
        tw[off + k] = make_float2((float)re, (float)im);
    };
    for (int k = 0; k < n1; ++k) put(p->tw_n1, k, n1);
    for (int k = 0; k < n2; ++k) put(p->tw_n2, k, n2);
    for (int k = 0; k < nlo; ++k) put(p->tw_lo, k, (long)L);
    for (int k = 0; k < nhi; ++k) {
        double re, im;
        unit_root((long)k << p->tw_s, (long)L, &re, &im);
        tw[p->tw_hi + k] = make_float2((float)re, (float)im);
    }
    if (p->acc_q > 1) {
        for (int k = 0; k < 256; ++k) put(p->tw_sub_lo, k, 1l << 16);
        for (int k = 0; k < 256; ++k) {
            double re, im;
            unit_root((long)k << 8, 1l << 16, &re, &im);
            tw[p->tw_sub_hi + k] = make_float2((float)re, (float)im);
        }
        for (int k = 0; k < 128; ++k) put(p->tw_sub_128, k, 128);
    }

    e = dalloc(p, &p->tw, ntw * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(p->tw, tw.data(), ntw * sizeof(float2), hipMemcpyHostToDevice);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = dalloc(p, &p->in_dev[i], B * M * p->in_bytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->in_ev[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->in_done[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->out_ready[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->d2h_u16[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = dalloc(p, &p->hist, M * sizeof(float2));
    if (e == hipSuccess) e = dalloc(p, &p->ring, (size_t)p->FBS * L * sizeof(float2));
    if (e == hipSuccess) e = dalloc(p, &p->H, D * L * sizeof(float2));
    if (e == hipSuccess) e = dalloc(p, &p->delay, D * L * sizeof(uint16_t));
    if (e == hipSuccess) e = dalloc(p, &p->scratch, (size_t)p->G * L * sizeof(float2));
    if (e == hipSuccess && p->accum_ok && p->acc_q > 1) {
        // the sub-unit spectra of one push: [D][B][acc_q][L / acc_q] complex64
        e = dalloc(p, &p->xpre, D * B * L * sizeof(float2));
        if (e == hipSuccess) e = dalloc(p, &p->sub_sync, D * B * sizeof(unsigned int));
    } else if (e == hipSuccess && p->accum_ok) {
        e = dalloc(p, &p->delay_acc, D * L);
        if (e == hipSuccess) e = hipMemset(p->delay_acc, 0, D * L);
        if (e == hipSuccess && p->perm_rl) e = dalloc(p, &p->ring_p, (size_t)p->FBS * L * sizeof(float2));
        if (e == hipSuccess && p->perm_rl) e = hipMemset(p->ring_p, 0, (size_t)p->FBS * L * sizeof(float2));
        if (e == hipSuccess) e = dalloc(p, &p->H_p, D * L * sizeof(float2));
        if (e == hipSuccess) e = hipMemset(p->H_p, 0, D * L * sizeof(float2));
    }
    if (e == hipSuccess && (p->out_flags & MSOSM_OUT_C64)) e = dalloc(p, &p->out_c64, D * B * M * sizeof(float2));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->d2h_c64, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->d2h_f16, hipEventDisableTiming);
    // copy streams at the default priority, non-blocking like the compute stream
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->in_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->out_stream, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess && (p->out_flags & MSOSM_OUT_U16); ++i)
        e = dalloc(p, &p->out_u16b[i], D * B * M * sizeof(uint32_t));
    p->out_u16 = p->out_u16b[0];
    if (e == hipSuccess && (p->out_flags & MSOSM_OUT_F16)) e = dalloc(p, &p->out_f16, D * B * M * sizeof(uint32_t));
    if (e == hipSuccess && (p->out_flags & MSOSM_OUT_F16)) e = dalloc(p, &p->out_scale, D * B * sizeof(float));
    if (e == hipSuccess) e = hipMemset(p->hist, 0, M * sizeof(float2));
    if (e == hipSuccess) e = hipMemset(p->ring, 0, (size_t)p->FBS * L * sizeof(float2));
    if (e == hipSuccess) e = hipMemset(p->H, 0, D * L * sizeof(float2));
    if (e == hipSuccess) e = hipMemset(p->delay, 0, D * L * sizeof(uint16_t));
    if (e != hipSuccess)
        return fail(MSOSM_ERR_ALLOC, std::string("plan_create: device allocation: ") + hipGetErrorString(e));
    p->dm_set.assign(D, 0);
    *out = p;
    return MSOSM_OK;
}

extern "C" int msosm_plan_destroy(msosm_plan *p)
{
    plan_free(p);
    return MSOSM_OK;
}

// copy L complex values to the device multiplied by 1/L (exact for normal values: a power of two), in
// natural order and, for a plan with the accumulate launch of one unit, also in its chirp layout (bin k
// at acc_chirp_pos(perm_pos(k)))
static hipError_t upload_scaled(const msosm_plan *p, float2 *dst, float2 *dst_p, const float *src)
{
    const int L = p->L;
    std::vector<float2> tmp((size_t)L);
    const float s = std::ldexp(1.0f, -ilog2(L));
    for (int k = 0; k < L; ++k) tmp[k] = make_float2(src[2 * k] * s, src[2 * k + 1] * s);
    hipError_t e = hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess || !dst_p) return e;
    std::vector<float2> perm((size_t)L);
    for (int k = 0; k < L; ++k) perm[acc_chirp_pos(perm_pos((uint32_t)k, p->perm_rl, p->perm_sl))] = tmp[k];
    return hipMemcpy(dst_p, perm.data(), perm.size() * sizeof(float2), hipMemcpyHostToDevice);
}

extern "C" int msosm_plan_set_dm(msosm_plan *p, int i, const int *delay_points, const float *H)
{
    if (!p || !delay_points || !H) return set_error(MSOSM_ERR_ARG, "plan_set_dm: null argument");
    if (i < 0 || i >= p->D) return set_error(MSOSM_ERR_ARG, "plan_set_dm: dm index out of range");
    std::vector<uint16_t> dl(p->L);
    for (int k = 0; k < p->L; ++k) {
        const int v = delay_points[k];
        if (v < 0 || v >= p->dc)
            return set_error(MSOSM_ERR_ARG, "plan_set_dm: delay_points[" + std::to_string(k) + "]=" + std::to_string(v) +
                                                " outside [0, delaycount=" + std::to_string(p->dc) + ")");
        dl[k] = (uint16_t)v;
    }
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipMemcpy(p->delay + (size_t)i * p->L, dl.data(), p->L * sizeof(uint16_t), hipMemcpyHostToDevice));
    // the device copy carries the inverse transform's 1/L (exact: a power of two)
    HIPCHK(upload_scaled(p, p->H + (size_t)i * p->L, p->H_p ? p->H_p + (size_t)i * p->L : nullptr, H));
    if (p->delay_acc) {
        // byte delays in thread order: position r S + p + 1024 t (perm_pos) goes to byte r S + 16 p + t
        std::vector<uint8_t> db((size_t)p->L);
        const uint32_t S = 1u << 14;
        for (int k = 0; k < p->L; ++k) {
            const uint32_t pos = perm_pos((uint32_t)k, p->perm_rl, p->perm_sl);
            const uint32_t r = pos / S, q = pos % S;
            db[(size_t)r * S + 16 * (q & 1023u) + (q >> 10)] = (uint8_t)delay_points[k];
        }
        HIPCHK(hipMemcpy(p->delay_acc + (size_t)i * p->L, db.data(), db.size(), hipMemcpyHostToDevice));
    }
    p->dm_set[i] = 1;
    return MSOSM_OK;
}

extern "C" int msosm_plan_set_dm_tangent(msosm_plan *p, int i, const float *dH)
{
    if (!p || !dH) return set_error(MSOSM_ERR_ARG, "plan_set_dm_tangent: null argument");
    if (i < 0 || i >= p->D) return set_error(MSOSM_ERR_ARG, "plan_set_dm_tangent: dm index out of range");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamSynchronize(p->stream));
    if (!p->Hdot) {
        const size_t hb = (size_t)p->D * p->L * sizeof(float2), yb = (size_t)p->D * p->B * p->M * sizeof(float2);
        hipError_t e = dalloc(p, &p->Hdot, hb);
        if (e == hipSuccess) e = dalloc(p, &p->tan_y, yb);
        if (e == hipSuccess) e = hipMemset(p->Hdot, 0, hb);
        // the accumulate layout of dH/dDM where a tangent push can run the accumulate form (not the
        // sub-unit form of 2^17 and 2^18, whose tangent pushes run the two-pass form)
        const bool acc_tan = p->accum_ok && p->acc_q == 1;
        if (e == hipSuccess && acc_tan) e = dalloc(p, &p->Hdot_p, hb);
        if (e == hipSuccess && acc_tan) e = hipMemset(p->Hdot_p, 0, hb);
        if (e != hipSuccess)
            return set_error(MSOSM_ERR_ALLOC, std::string("plan_set_dm_tangent: device allocation: ") + hipGetErrorString(e));
        p->tan_set.assign(p->D, 0);
    }
    HIPCHK(upload_scaled(p, p->Hdot + (size_t)i * p->L, p->Hdot_p ? p->Hdot_p + (size_t)i * p->L : nullptr, dH));
    p->tan_set[i] = 1;
    return MSOSM_OK;
}

static int record_begin(msosm_plan *p, EventPair &ev)
{
    if (p->pool.empty()) {
        EventPair e;
        HIPCHK(hipEventCreate(&e.a));
        HIPCHK(hipEventCreate(&e.b));
        p->pool.push_back(e);
    }
    ev = p->pool.back();
    p->pool.pop_back();
    HIPCHK(hipEventRecord(ev.a, p->stream));
    return MSOSM_OK;
}

static int collect_profile(msosm_plan *p)
{
    for (int k = 0; k < 2; ++k) {
        for (auto &e : p->pending[k]) {
            HIPCHK(hipEventSynchronize(e.b));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
            p->prof_ms[k] += ms;
            p->prof_n[k] += 1;
            p->pool.push_back(e);
        }
        p->pending[k].clear();
    }
    return MSOSM_OK;
}

// Arguments of every launch of one push (no launch): forward transforms of this push's blocks into
// the ring, the overlap history for the next push, the fold bins (if a fold is attached) and the
// dedispersion with the attached fold.
void msosm::plan_push_args(msosm_plan *p, const void *src, PushArgs *out, bool allow_sub)
{
    PushArgs &a = *out;
    a = PushArgs{};
    const float2 *tw = p->tw;
    FwdArgs &fa = a.fwd;
    fa.in = src;
    fa.in_fmt = p->in_fmt;
    fa.hist = p->hist;
    fa.ring = p->ring;
    fa.FBS = p->FBS;
    fa.g0 = p->g0;
    fa.B = p->B;
    fa.scratch = p->scratch;
    fa.tw_n1 = tw + p->tw_n1;
    fa.tw_n2 = tw + p->tw_n2;
    fa.tw_lo = tw + p->tw_lo;
    fa.tw_hi = tw + p->tw_hi;
    fa.tw_s = p->tw_s;
    fa.ring_p = p->ring_p; // a second copy only where the accumulate layout differs
    fa.perm_rl = p->perm_rl;
    fa.perm_sl = p->perm_sl;
    a.log2L = p->log2L;
    a.fwd_grid = std::min(p->B, p->G);
    a.hist = HistArgs{p->hist, src, p->in_fmt, (long long)(p->B - 1) * p->M, p->M};
    InvArgs &ia = a.inv;
    ia.ring = p->ring;
    ia.FBS = p->FBS;
    ia.g0 = p->g0;
    ia.B = p->B;
    ia.D = p->D;
    ia.H = p->H;
    ia.delay = p->delay;
    ia.scratch = p->scratch;
    ia.out_c64 = p->out_c64;
    ia.out_u16 = p->out_u16;
    ia.out_f16 = p->out_f16;
    ia.out_scale = p->out_scale;
    ia.tw_n1 = fa.tw_n1;
    ia.tw_n2 = fa.tw_n2;
    ia.tw_lo = fa.tw_lo;
    ia.tw_hi = fa.tw_hi;
    ia.tw_s = p->tw_s;
    ia.order = p->order;
    ia.ring_p = p->ring_p ? p->ring_p : p->ring;
    ia.H_p = p->H_p;
    ia.delay_acc = p->delay_acc;
    // the sub-unit accumulate launch where the schedule asks for it and the caller runs it (not for the
    // one-launch uneven kernel or a tangent push): its tables, and its order of the fold bins
    const bool sub = allow_sub && p->acc_q > 1 && (p->order & MSOSM_SCHEDULE_ACCUM);
    if (p->acc_q > 1 && !sub) ia.order &= ~MSOSM_SCHEDULE_ACCUM;
    if (sub) {
        ia.sub_lo = tw + p->tw_sub_lo;
        ia.sub_hi = tw + p->tw_sub_hi;
        ia.sub_128 = tw + p->tw_sub_128;
        ia.sub_s = p->tw_sub_s;
        ia.xpre = p->xpre;
        ia.sub_sync = p->sub_sync;
    }
    if (p->fold) {
        msosm_fold *f = p->fold;
        a.has_bins = true;
        a.bins = BinsArgs{p->bins, (long long)p->B * p->M, f->nbins, f->t_start, f->dt, f->period, sub ? p->acc_q : 1,
                          p->M};
        ia.bins = p->bins;
        ia.fold_sum = f->sums;
        ia.fold_cnt = f->cnts;
        ia.nbins = (int)f->nbins;
    }
}

// The part of a push before the dedispersion launch: launches the forward transforms, the history
// copy and the fold bins, and returns the dedispersion arguments (the launch is the caller's).
int msosm::plan_push_prepare(msosm_plan *p, const void *src, InvArgs *out, bool allow_sub)
{
    PushArgs a;
    plan_push_args(p, src, &a, allow_sub);
    EventPair ev{};
    if (p->profile && record_begin(p, ev) != MSOSM_OK) return MSOSM_ERR_HIP;
    HIPCHK(launch_fwd(a.log2L, a.fwd, a.fwd_grid, p->G, p->stream));
    if (p->profile) {
        HIPCHK(hipEventRecord(ev.b, p->stream));
        p->pending[0].push_back(ev);
    }
    HIPCHK(launch_hist(a.hist.hist, a.hist.in, a.hist.fmt, a.hist.first, a.hist.M, p->stream));
    if (a.has_bins && a.bins.q > 1)
        HIPCHK(launch_bins_sub(a.bins.bins, a.bins.n, a.bins.nbins, a.bins.t_start, a.bins.dt, a.bins.period, a.bins.q,
                               a.bins.m, p->stream));
    else if (a.has_bins)
        HIPCHK(launch_bins(a.bins.bins, a.bins.n, a.bins.nbins, a.bins.t_start, a.bins.dt, a.bins.period, p->stream));
    *out = a.inv;
    return MSOSM_OK;
}

// The part after the dedispersion launch: advance the ring position and the fold clock.
void msosm::plan_push_commit(msosm_plan *p)
{
    const size_t nsamp = (size_t)p->B * p->M;
    p->g0 += p->B;
    p->pushes += 1;
    if (p->fold) p->fold->t_start += static_cast<double>(nsamp) * p->fold->dt; // fold_gpu.cpp:53
}

int msosm::plan_check_ready(const msosm_plan *p)
{
    for (int i = 0; i < p->D; ++i)
        if (!p->dm_set[i]) return set_error(MSOSM_ERR_STATE, "plan_push: DM " + std::to_string(i) + " has no design");
    if (!p->fold && !p->out_flags)
        return set_error(MSOSM_ERR_STATE, "plan_push: no output selected and no fold attached");
    return MSOSM_OK;
}

int msosm::plan_log2L(const msosm_plan *p) { return p->log2L; }
int msosm::plan_workgroups(const msosm_plan *p) { return p->G; }
int msosm::plan_num_units(const msosm_plan *p) { return p->D * p->B; }
bool msosm::plan_ring32(const msosm_plan *p) { return (unsigned long long)p->FBS * p->L * sizeof(float2) < (1ull << 32); }
int msosm::plan_launch_mode(const msosm_plan *p) { return p->launch_mode; }
hipStream_t msosm::plan_stream(const msosm_plan *p) { return p->stream; }

// One push: stage the input, forward transforms into the ring, then the dedispersion launch(es).
// With tfold, the dedispersion runs twice: with H into the primal buffer (folding |y|^2 into the
// attached fold, if any), then with dH/dDM, folding 2 Re(conj(y) dy/dDM) into tfold.
static int push_impl(msosm_plan *p, const void *input, int mode, msosm_fold *tfold)
{
    if (!p || !input) return set_error(MSOSM_ERR_ARG, "plan_push: null argument");
    if (mode != MSOSM_INPUT_HOST_ASYNC && mode != MSOSM_INPUT_DEVICE && mode != MSOSM_INPUT_HOST_SYNC)
        return set_error(MSOSM_ERR_ARG, "plan_push: unknown input mode");
    for (int i = 0; i < p->D; ++i)
        if (!p->dm_set[i]) return set_error(MSOSM_ERR_STATE, "plan_push: DM " + std::to_string(i) + " has no design");
    if (!tfold && !p->fold && !p->out_flags)
        return set_error(MSOSM_ERR_STATE, "plan_push: no output selected and no fold attached");
    if (tfold) {
        if (!p->Hdot) return set_error(MSOSM_ERR_STATE, "plan_push_tangent: no dH/dDM uploaded");
        for (int i = 0; i < p->D; ++i)
            if (!p->tan_set[i])
                return set_error(MSOSM_ERR_STATE, "plan_push_tangent: DM " + std::to_string(i) + " has no dH/dDM");
        if (tfold->S != p->D || tfold->device != p->device)
            return set_error(MSOSM_ERR_ARG, "plan_push_tangent: tangent fold needs one series per DM on the plan's device");
        if (tfold->nbins > (unsigned long)p->max_fold_bins)
            return set_error(MSOSM_ERR_UNSUPPORTED, "plan_push_tangent: at most " + std::to_string(p->max_fold_bins) +
                                                        " bins in LDS");
        if (tfold->size != (unsigned long)p->B * p->M)
            return set_error(MSOSM_ERR_ARG, "plan_push_tangent: tangent fold size must be batch*M");
        if (p->fold && (p->fold->nbins != tfold->nbins || p->fold->period != tfold->period || p->fold->dt != tfold->dt ||
                        p->fold->t_start != tfold->t_start))
            return set_error(MSOSM_ERR_ARG, "plan_push_tangent: the attached fold and the tangent fold disagree on bins");
    }
    HIPCHK(hipSetDevice(p->device));
    const size_t nsamp = (size_t)p->B * p->M;
    const void *src = input;
    const int slot = (int)(p->pushes & 1);
    if (mode == MSOSM_INPUT_HOST_SYNC) {
        // the staging buffer was last read by the push two calls ago; wait for that to finish
        HIPCHK(hipEventSynchronize(p->in_ev[slot]));
        HIPCHK(hipMemcpy(p->in_dev[slot], input, nsamp * p->in_bytes, hipMemcpyHostToDevice));
        src = p->in_dev[slot];
    } else if (mode == MSOSM_INPUT_HOST_ASYNC) {
        // on the copy-in stream, once the push two calls ago has read the slot: the copy runs under the
        // previous push's kernels, and this push's transforms wait for it
        HIPCHK(hipStreamWaitEvent(p->in_stream, p->in_ev[slot], 0));
        HIPCHK(hipMemcpyAsync(p->in_dev[slot], input, nsamp * p->in_bytes, hipMemcpyHostToDevice, p->in_stream));
        HIPCHK(hipEventRecord(p->in_done[slot], p->in_stream));
        HIPCHK(hipStreamWaitEvent(p->stream, p->in_done[slot], 0));
        src = p->in_dev[slot];
    }
    // this push's uint16 buffer (parity), free once its previous copy-out (two pushes ago) is done; the
    // single complex64 and FP16 buffers once their last copy-out is done
    if (p->out_u16b[0]) {
        p->out_u16 = p->out_u16b[p->pushes & 1];
        HIPCHK(hipStreamWaitEvent(p->stream, p->d2h_u16[p->pushes & 1], 0));
    }
    if (p->out_c64) HIPCHK(hipStreamWaitEvent(p->stream, p->d2h_c64, 0));
    if (p->out_f16) HIPCHK(hipStreamWaitEvent(p->stream, p->d2h_f16, 0));
    InvArgs ia{};
    if (plan_push_prepare(p, src, &ia, tfold == nullptr) != MSOSM_OK) return MSOSM_ERR_HIP;
    if (mode != MSOSM_INPUT_DEVICE) HIPCHK(hipEventRecord(p->in_ev[slot], p->stream));
    if (tfold && !p->fold)
        HIPCHK(launch_bins(p->bins, (long long)nsamp, tfold->nbins, tfold->t_start, tfold->dt, tfold->period,
                           p->stream));
    EventPair ev{};
    auto run_inv = [&](const InvArgs &x) -> int {
        if (p->launch_mode == MSOSM_LAUNCH_PER_DM) {
            // the plain form: one launch per DM over that DM's blocks
            const size_t ser = (size_t)p->B * p->M;
            for (int d = 0; d < p->D; ++d) {
                InvArgs one = x;
                one.D = 1;
                one.H = x.H + (size_t)d * p->L;
                one.delay = x.delay + (size_t)d * p->L;
                one.H_p = x.H_p + (size_t)d * p->L;
                if (x.xpre) one.xpre = x.xpre + (size_t)d * p->B * p->L;
                if (x.sub_sync) one.sub_sync = x.sub_sync + (size_t)d * p->B;
                if (x.delay_acc) one.delay_acc = x.delay_acc + (size_t)d * p->L;
                if (x.out_c64) one.out_c64 = x.out_c64 + d * ser;
                if (x.out_u16) one.out_u16 = x.out_u16 + d * ser;
                if (x.out_f16) one.out_f16 = x.out_f16 + d * ser;
                if (x.out_scale) one.out_scale = x.out_scale + (size_t)d * p->B;
                if (x.tangent_ref) one.tangent_ref = x.tangent_ref + d * ser;
                if (x.fold_sum) {
                    one.fold_sum = x.fold_sum + (size_t)d * x.nbins;
                    one.fold_cnt = x.fold_cnt + (size_t)d * x.nbins;
                }
                HIPCHK(launch_inv(p->log2L, one, std::min(p->B, p->G), p->stream));
            }
        } else {
            HIPCHK(launch_inv(p->log2L, x, p->sched_wg ? p->sched_wg : p->G, p->stream));
        }
        return MSOSM_OK;
    };
    if (p->profile && record_begin(p, ev) != MSOSM_OK) return MSOSM_ERR_HIP;
    if (!tfold) {
        if (run_inv(ia) != MSOSM_OK) return MSOSM_ERR_HIP;
    } else {
        // primal series into the pairing buffer (no user outputs in a tangent push), then the tangent
        InvArgs prim = ia;
        prim.out_c64 = p->tan_y;
        prim.out_u16 = nullptr;
        prim.out_f16 = nullptr;
        prim.out_scale = nullptr;
        if (run_inv(prim) != MSOSM_OK) return MSOSM_ERR_HIP;
        InvArgs tan = prim;
        tan.H = p->Hdot;
        tan.H_p = p->Hdot_p;
        tan.out_c64 = nullptr;
        tan.tangent_ref = p->tan_y;
        tan.bins = p->bins;
        tan.fold_sum = tfold->sums;
        tan.fold_cnt = tfold->cnts;
        tan.nbins = (int)tfold->nbins;
        if (run_inv(tan) != MSOSM_OK) return MSOSM_ERR_HIP;
    }
    if (p->profile) {
        HIPCHK(hipEventRecord(ev.b, p->stream));
        p->pending[1].push_back(ev);
    }
    HIPCHK(hipEventRecord(p->out_ready[p->pushes & 1], p->stream));
    plan_push_commit(p);
    if (tfold) tfold->t_start += static_cast<double>(nsamp) * tfold->dt;
    return MSOSM_OK;
}

extern "C" int msosm_plan_push(msosm_plan *p, const void *input, int mode) { return push_impl(p, input, mode, nullptr); }

extern "C" int msosm_plan_push_tangent(msosm_plan *p, const void *input, int mode, msosm_fold *tangent_fold)
{
    if (!p || !tangent_fold) return set_error(MSOSM_ERR_ARG, "plan_push_tangent: null argument");
    if (!p->bins) {
        HIPCHK(hipSetDevice(p->device));
        hipError_t e = dalloc(p, &p->bins, (size_t)p->B * p->M * sizeof(uint16_t));
        if (e != hipSuccess) return set_error(MSOSM_ERR_ALLOC, std::string("plan_push_tangent: ") + hipGetErrorString(e));
    }
    return push_impl(p, input, mode, tangent_fold);
}

extern "C" int msosm_plan_output_c64(msosm_plan *p, void **ptr)
{
    if (!p || !ptr) return set_error(MSOSM_ERR_ARG, "plan_output_c64: null argument");
    if (!p->out_c64) return set_error(MSOSM_ERR_STATE, "plan_output_c64: plan has no complex64 output");
    *ptr = p->out_c64;
    return MSOSM_OK;
}

extern "C" int msosm_plan_output_u16(msosm_plan *p, void **ptr)
{
    if (!p || !ptr) return set_error(MSOSM_ERR_ARG, "plan_output_u16: null argument");
    if (!p->out_u16b[0]) return set_error(MSOSM_ERR_STATE, "plan_output_u16: plan has no uint16 output");
    *ptr = p->out_u16b[p->pushes ? (p->pushes - 1) & 1 : 0]; // the last push's buffer
    return MSOSM_OK;
}

extern "C" int msosm_plan_output_f16(msosm_plan *p, void **data, void **scales)
{
    if (!p || !data || !scales) return set_error(MSOSM_ERR_ARG, "plan_output_f16: null argument");
    if (!p->out_f16) return set_error(MSOSM_ERR_STATE, "plan_output_f16: plan has no FP16 output");
    *data = p->out_f16;
    *scales = p->out_scale;
    return MSOSM_OK;
}

extern "C" int msosm_plan_get_output(msosm_plan *p, int which, void *host)
{
    if (!p || !host) return set_error(MSOSM_ERR_ARG, "plan_get_output: null argument");
    if (p->pushes == 0) return set_error(MSOSM_ERR_STATE, "plan_get_output: nothing pushed yet");
    const size_t n = (size_t)p->D * p->B * p->M;
    HIPCHK(hipSetDevice(p->device));
    // the last push's outputs, on the copy-out stream (get_output, msosm_gpu_batch.cpp:238-245)
    const int last = (int)((p->pushes - 1) & 1);
    hipStream_t os = p->out_stream;
    if (which == MSOSM_OUT_C64) {
        if (!p->out_c64) return set_error(MSOSM_ERR_STATE, "plan_get_output: plan has no complex64 output");
        HIPCHK(hipStreamWaitEvent(os, p->out_ready[last], 0));
        HIPCHK(hipMemcpyAsync(host, p->out_c64, n * sizeof(float2), hipMemcpyDeviceToHost, os));
        HIPCHK(hipEventRecord(p->d2h_c64, os));
    } else if (which == MSOSM_OUT_U16) {
        if (!p->out_u16b[0]) return set_error(MSOSM_ERR_STATE, "plan_get_output: plan has no uint16 output");
        HIPCHK(hipStreamWaitEvent(os, p->out_ready[last], 0));
        HIPCHK(hipMemcpyAsync(host, p->out_u16b[last], n * sizeof(uint32_t), hipMemcpyDeviceToHost, os));
        HIPCHK(hipEventRecord(p->d2h_u16[last], os));
    } else if (which == MSOSM_OUT_F16) {
        if (!p->out_f16) return set_error(MSOSM_ERR_STATE, "plan_get_output: plan has no FP16 output");
        HIPCHK(hipStreamWaitEvent(os, p->out_ready[last], 0));
        HIPCHK(hipMemcpyAsync(host, p->out_f16, n * sizeof(uint32_t), hipMemcpyDeviceToHost, os));
        HIPCHK(hipMemcpyAsync(static_cast<char *>(host) + n * sizeof(uint32_t), p->out_scale,
                              (size_t)p->D * p->B * sizeof(float), hipMemcpyDeviceToHost, os));
        HIPCHK(hipEventRecord(p->d2h_f16, os));
    } else {
        return set_error(MSOSM_ERR_ARG, "plan_get_output: which must be MSOSM_OUT_C64, MSOSM_OUT_U16 or MSOSM_OUT_F16");
    }
    return MSOSM_OK;
}

extern "C" int msosm_plan_get_output_stream(msosm_plan *p, void **stream)
{
    if (!p || !stream) return set_error(MSOSM_ERR_ARG, "plan_get_output_stream: null argument");
    *stream = p->out_stream;
    return MSOSM_OK;
}

extern "C" int msosm_plan_synchronize(msosm_plan *p)
{
    if (!p) return set_error(MSOSM_ERR_ARG, "plan_synchronize: null plan");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamSynchronize(p->in_stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipStreamSynchronize(p->out_stream));
    return MSOSM_OK;
}

extern "C" int msosm_plan_reset(msosm_plan *p)
{
    if (!p) return set_error(MSOSM_ERR_ARG, "plan_reset: null plan");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(launch_fill(p->ring, (long long)p->FBS * p->L * sizeof(float2), 0u, p->stream));
    if (p->ring_p) HIPCHK(launch_fill(p->ring_p, (long long)p->FBS * p->L * sizeof(float2), 0u, p->stream));
    HIPCHK(launch_fill(p->hist, (long long)p->M * sizeof(float2), 0u, p->stream));
    p->g0 = 0;
    return MSOSM_OK;
}

extern "C" int msosm_plan_get_stream(msosm_plan *p, void **stream)
{
    if (!p || !stream) return set_error(MSOSM_ERR_ARG, "plan_get_stream: null argument");
    *stream = p->stream;
    return MSOSM_OK;
}

extern "C" int msosm_plan_get_info(msosm_plan *p, msosm_plan_info *info)
{
    if (!p || !info) return set_error(MSOSM_ERR_ARG, "plan_get_info: null argument");
    info->L = p->L;
    info->FBS = p->FBS;
    info->workgroups = p->G;
    info->blocks_pushed = p->g0;
    info->device_bytes = p->device_bytes;
    return MSOSM_OK;
}

extern "C" int msosm_plan_profile(msosm_plan *p, int enable)
{
    if (!p) return set_error(MSOSM_ERR_ARG, "plan_profile: null plan");
    if (collect_profile(p) != MSOSM_OK) return MSOSM_ERR_HIP;
    p->profile = enable != 0;
    p->prof_ms[0] = p->prof_ms[1] = 0;
    p->prof_n[0] = p->prof_n[1] = 0;
    return MSOSM_OK;
}

extern "C" int msosm_plan_profile_read(msosm_plan *p, int kernel, double *total_ms, long long *launches)
{
    if (!p || kernel < 0 || kernel > 1) return set_error(MSOSM_ERR_ARG, "plan_profile_read: bad arguments");
    HIPCHK(hipSetDevice(p->device));
    if (collect_profile(p) != MSOSM_OK) return MSOSM_ERR_HIP;
    if (total_ms) *total_ms = p->prof_ms[kernel];
    if (launches) *launches = p->prof_n[kernel];
    return MSOSM_OK;
}

extern "C" int msosm_plan_set_schedule(msosm_plan *p, int workgroups, int order)
{
    if (!p) return set_error(MSOSM_ERR_ARG, "plan_set_schedule: null plan");
    if (workgroups < 0 || workgroups > p->G)
        return set_error(MSOSM_ERR_ARG, "plan_set_schedule: workgroups must be in [0, " + std::to_string(p->G) + "]");
    if ((order & ~(MSOSM_ORDER_MASK | MSOSM_SCHEDULE_SPLIT | MSOSM_SCHEDULE_ACCUM)) ||
        (order & MSOSM_ORDER_MASK) > MSOSM_ORDER_XCD_TILE2)
        return set_error(MSOSM_ERR_ARG, "plan_set_schedule: unknown order");
    p->sched_wg = workgroups;
    p->order = order;
    return MSOSM_OK;
}

extern "C" int msosm_plan_get_schedule(msosm_plan *p, int *workgroups, int *order)
{
    if (!p) return set_error(MSOSM_ERR_ARG, "plan_get_schedule: null plan");
    if (workgroups) *workgroups = p->sched_wg ? p->sched_wg : p->G;
    if (order) *order = p->order;
    return MSOSM_OK;
}

extern "C" int msosm_plan_autotune(msosm_plan *p, const void *input, int pushes, int *best_wg, int *best_order,
                                   double *best_ms)
{
    if (!p || !input) return set_error(MSOSM_ERR_ARG, "plan_autotune: null argument");
    if (pushes < 2) return set_error(MSOSM_ERR_ARG, "plan_autotune: need at least 2 pushes of input");
    if (p->launch_mode != MSOSM_LAUNCH_PERSISTENT)
        return set_error(MSOSM_ERR_STATE, "plan_autotune: the schedule applies to the persistent launch only");
    HIPCHK(hipSetDevice(p->device));
    const bool was_profiling = p->profile;
    if (collect_profile(p) != MSOSM_OK) return MSOSM_ERR_HIP;
    const size_t step = (size_t)p->B * p->M * p->in_bytes;
    const int half = std::max(8, (p->G / 2 / 8) * 8);
    const int wgs[2] = {p->G, half};
    double best = 1e300;
    int bw = p->G, bo = MSOSM_ORDER_XCD_DM_MAJOR;
    // launch forms: sequential, role split (FFT length, plain epilogue) and accumulate (FFT length)
    std::vector<int> forms{0};
    if (split_supported(p->log2L) && !(p->out_flags & MSOSM_OUT_F16)) forms.push_back(MSOSM_SCHEDULE_SPLIT);
    if (p->accum_ok) forms.push_back(MSOSM_SCHEDULE_ACCUM);
    const int norders = MSOSM_ORDER_XCD_TILE2 + 1;
    for (int wi = 0; wi < (half < p->G ? 2 : 1); ++wi)
        for (int c = 0; c < norders * (int)forms.size(); ++c) {
            const int o = (c % norders) | forms[c / norders];
            p->sched_wg = wgs[wi];
            p->order = o;
            int rc = msosm_plan_reset(p);
            if (rc == MSOSM_OK && p->fold) rc = msosm_fold_reset(p->fold);
            if (rc == MSOSM_OK) rc = push_impl(p, input, MSOSM_INPUT_DEVICE, nullptr);
            if (rc == MSOSM_OK) rc = collect_profile(p);
            if (rc != MSOSM_OK) return rc;
            p->profile = true;
            p->prof_ms[1] = 0;
            p->prof_n[1] = 0;
            for (int k = 1; k < pushes && rc == MSOSM_OK; ++k)
                rc = push_impl(p, static_cast<const char *>(input) + k * step, MSOSM_INPUT_DEVICE, nullptr);
            if (rc == MSOSM_OK) rc = collect_profile(p);
            p->profile = false;
            if (rc != MSOSM_OK) return rc;
            const double ms = p->prof_ms[1] / (double)std::max(1LL, p->prof_n[1]);
            if (ms < best) {
                best = ms;
                bw = wgs[wi];
                bo = o;
            }
        }
    p->sched_wg = bw == p->G ? 0 : bw;
    p->order = bo;
    int rc = msosm_plan_reset(p);
    if (rc == MSOSM_OK && p->fold) rc = msosm_fold_reset(p->fold);
    if (rc != MSOSM_OK) return rc;
    p->profile = was_profiling;
    p->prof_ms[0] = p->prof_ms[1] = 0;
    p->prof_n[0] = p->prof_n[1] = 0;
    if (best_wg) *best_wg = bw;
    if (best_order) *best_order = bo;
    if (best_ms) *best_ms = best;
    return MSOSM_OK;
}

extern "C" int msosm_plan_attach_fold(msosm_plan *p, msosm_fold *f)
{
    if (!p) return set_error(MSOSM_ERR_ARG, "plan_attach_fold: null plan");
    HIPCHK(hipSetDevice(p->device));
    if (!f) {
        p->fold = nullptr;
        return MSOSM_OK;
    }
    const size_t nsamp = (size_t)p->B * p->M;
    if (f->size != nsamp)
        return set_error(MSOSM_ERR_ARG, "plan_attach_fold: fold size " + std::to_string(f->size) +
                                            " != batch*M = " + std::to_string(nsamp));
    if (f->S != p->D) return set_error(MSOSM_ERR_ARG, "plan_attach_fold: fold needs one series per DM");
    if (f->nbins > (unsigned long)p->max_fold_bins)
        return set_error(MSOSM_ERR_UNSUPPORTED, "plan_attach_fold: the fused fold at this FFT length holds at most " +
                                                    std::to_string(p->max_fold_bins) + " bins in LDS");
    if (f->device != p->device) return set_error(MSOSM_ERR_ARG, "plan_attach_fold: fold and plan on different devices");
    if (!p->bins) {
        hipError_t e = dalloc(p, &p->bins, nsamp * sizeof(uint16_t));
        if (e != hipSuccess) return set_error(MSOSM_ERR_ALLOC, std::string("plan_attach_fold: ") + hipGetErrorString(e));
    }
    p->fold = f;
    return MSOSM_OK;
}

extern "C" int msosm_plan_set_launch_mode(msosm_plan *p, int mode)
{
    if (!p || (mode != MSOSM_LAUNCH_PERSISTENT && mode != MSOSM_LAUNCH_PER_DM))
        return set_error(MSOSM_ERR_ARG, "plan_set_launch_mode: unknown mode");
    p->launch_mode = mode;
    return MSOSM_OK;
}
