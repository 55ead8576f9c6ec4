// C ABI of the GPU pulsar simulator: SimulatedComplexSignal::generate_pulsar_signal_block
// (simulated_complex_signal.cpp:74-277) with the dispersion injected on the GPU. The host builds the
// noise periods and the chirp exactly as the reference does (sim_host.cpp); each chunk of 128 periods
// is then transformed, multiplied by the chirp, transformed back, scaled by 1/N and its central
// periods converted to the reference's offset-binary uint16 pairs, all on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "capi_internal.h"
#include "kernels.h"
#include "msosm_amd.h"
#include "sim_host.h"

using namespace msosm;

namespace {

// device buffers of one simulation, freed on every exit path
struct SimDev {
    std::vector<void *> ptrs;
    hipStream_t st = nullptr;
    ~SimDev()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void *p : ptrs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    hipError_t alloc(void **p, size_t bytes)
    {
        const hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// the engine's twiddle tables of length 2^lg (n1, n2, lo, hi), as the plans build them
struct SimTw {
    float2 *dev = nullptr;
    int n1 = 0, n2 = 0, lo = 0, hi = 0, s = 0;
};

hipError_t make_tw(SimDev &d, int lg, SimTw *t)
{
    int n1, n2;
    split_length(lg, &n1, &n2);
    const long L = 1L << lg;
    t->s = (lg + 1) / 2;
    const int nlo = 1 << t->s, nhi = (int)(L >> t->s);
    t->n1 = 0;
    t->n2 = n1;
    t->lo = n1 + n2;
    t->hi = n1 + n2 + nlo;
    std::vector<float2> tw(n1 + n2 + nlo + nhi);
    auto put = [&](int off, long k, long kk, long n) {
        double re, im;
        unit_root(kk, n, &re, &im);
        tw[off + k] = make_float2((float)re, (float)im);
    };
    for (int k = 0; k < n1; ++k) put(t->n1, k, k, n1);
    for (int k = 0; k < n2; ++k) put(t->n2, k, k, n2);
    for (int k = 0; k < nlo; ++k) put(t->lo, k, k, L);
    for (int k = 0; k < nhi; ++k) put(t->hi, k, (long)k << t->s, L);
    hipError_t e = d.alloc(reinterpret_cast<void **>(&t->dev), tw.size() * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(t->dev, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice);
    return e;
}

void set_tw(SimFftArgs &a, const SimTw &t)
{
    a.tw_n1 = t.dev + t.n1;
    a.tw_n2 = t.dev + t.n2;
    a.tw_lo = t.dev + t.lo;
    a.tw_hi = t.dev + t.hi;
    a.tw_s = t.s;
}

// one complex FFT of 2^lg points on the device: one workgroup up to 2^22 points, above that the
// six-step form N = A B over two in-workgroup lengths (n = c + B a, k = a' + A c': columns of length A
// at stride B times the twiddle W_N^{a' c}, then rows of length B written transposed in natural order)
struct SimPlan {
    int lg = 0, lga = 0, lgb = 0, grid = 1;
    bool one = true;
    long long N = 0, A = 0, B = 0;
    SimTw ta, tb;
    float2 *scr = nullptr, *tmp = nullptr;
};

bool sim_plan_supported(int lg)
{
    return sim_length_supported(lg) || (sim_length_supported((lg + 1) / 2) && sim_length_supported(lg / 2));
}

hipError_t make_plan(SimDev &d, int lg, int cus, SimPlan *p)
{
    p->lg = lg;
    p->one = sim_length_supported(lg);
    p->lga = (lg + 1) / 2;
    p->lgb = lg / 2;
    p->N = 1LL << lg;
    p->A = 1LL << p->lga;
    p->B = 1LL << p->lgb;
    p->grid = p->one ? 1 : std::max(8, cus);
    const long long Lsub = p->one ? p->N : p->A; // A >= B
    hipError_t e = d.alloc(reinterpret_cast<void **>(&p->scr), (size_t)p->grid * Lsub * sizeof(float2));
    if (e == hipSuccess && !p->one) e = d.alloc(reinterpret_cast<void **>(&p->tmp), p->N * sizeof(float2));
    if (e == hipSuccess) e = make_tw(d, p->one ? lg : p->lga, &p->ta);
    if (e == hipSuccess && !p->one) e = make_tw(d, p->lgb, &p->tb);
    return e;
}

// dst <- the transform of src (direction dir) through the output stage `mode`; `a` carries the
// stage's chirp or extraction fields. With mode SIM_EXTRACT, dst is unused.
hipError_t run_plan(const SimPlan &p, const float2 *src, float2 *dst, int dir, int mode, SimFftArgs a, hipStream_t st)
{
    a.scratch = p.scr;
    a.big_n = p.N;
    if (p.one) {
        set_tw(a, p.ta);
        a.ntrans = 1;
        a.in_t = a.out_t = 0;
        a.in_e = a.out_e = 1;
        a.src = src;
        a.dst = dst;
        a.mode = mode;
        return launch_sim_fft(p.lg, dir, a, p.grid, st);
    }
    set_tw(a, p.ta); // columns
    a.ntrans = p.B;
    a.in_t = a.out_t = 1;
    a.in_e = a.out_e = p.B;
    a.src = src;
    a.dst = p.tmp;
    a.mode = SIM_TWIDDLE;
    hipError_t e = launch_sim_fft(p.lga, dir, a, p.grid, st);
    if (e != hipSuccess) return e;
    set_tw(a, p.tb); // rows
    a.ntrans = p.A;
    a.in_t = p.B;
    a.in_e = 1;
    a.out_t = 1;
    a.out_e = p.A;
    a.src = p.tmp;
    a.dst = dst;
    a.mode = mode;
    return launch_sim_fft(p.lgb, dir, a, p.grid, st);
}

int log2_of(unsigned long n)
{
    int lg = 0;
    while ((1UL << lg) < n) ++lg;
    return lg;
}

int host_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

} // namespace

#define SIMCHK(x)                                                                                                      \
    do {                                                                                                               \
        const hipError_t e_ = (x);                                                                                     \
        if (e_ != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("simulate: ") + hipGetErrorString(e_));     \
    } while (0)

extern "C" int msosm_host_alloc(unsigned long long bytes, void **ptr)
{
    if (!ptr) return set_error(MSOSM_ERR_ARG, "host_alloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return MSOSM_OK;
    const hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return set_error(MSOSM_ERR_ALLOC, std::string("host_alloc: ") + hipGetErrorString(e));
    }
    return MSOSM_OK;
}

extern "C" int msosm_host_free(void *ptr)
{
    if (!ptr) return MSOSM_OK;
    const hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? MSOSM_OK : set_error(MSOSM_ERR_HIP, std::string("host_free: ") + hipGetErrorString(e));
}

extern "C" int msosm_simulate_params(float bw, float dm, float f0, float period, unsigned long dummy_repeat,
                                     unsigned long valid_repeat, unsigned long long *out)
{
    SimParams p;
    if (!out) return set_error(MSOSM_ERR_ARG, "simulate_params: null output");
    if (sim_params(bw, dm, f0, period, dummy_repeat, valid_repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_params: bw, f0, period > 0, dm >= 0 and at least one period");
    const unsigned long long v[MSOSM_SIM_NPARAMS] = {p.Np, p.Nd, p.range, p.fftpoint, p.signal_size, p.periods_len,
                                                     p.start, p.mul_repeat};
    std::copy(v, v + MSOSM_SIM_NPARAMS, out);
    return MSOSM_OK;
}

extern "C" int msosm_simulate_noise(float bw, float dm, float f0, float period, unsigned long dummy_repeat,
                                    unsigned long valid_repeat, float *periods, unsigned long long len)
{
    SimParams p;
    if (sim_params(bw, dm, f0, period, dummy_repeat, valid_repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_noise: bw, f0, period > 0, dm >= 0 and at least one period");
    if (!periods || len < p.periods_len) return set_error(MSOSM_ERR_ARG, "simulate_noise: buffer too small");
    sim_noise(p, periods);
    return MSOSM_OK;
}

extern "C" int msosm_simulate_chirp(float bw, float dm, float f0, unsigned long long fftpoint, float *H)
{
    if (!H || fftpoint == 0 || (fftpoint & (fftpoint - 1))) return set_error(MSOSM_ERR_ARG, "simulate_chirp: bad length");
    sim_chirp(bw, dm, f0, (unsigned long)fftpoint, H, host_threads());
    return MSOSM_OK;
}

extern "C" int msosm_simulate_pulsar_block(float bw, float dm, float f0, float period, unsigned long dummy_repeat,
                                           unsigned long valid_repeat, int device, uint32_t *out_u16, float *out_c64,
                                           unsigned long long len)
{
    SimParams p;
    if (sim_params(bw, dm, f0, period, dummy_repeat, valid_repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate: bw, f0, period > 0, dm >= 0 and at least one period");
    if ((!out_u16 && !out_c64) || len < p.signal_size)
        return set_error(MSOSM_ERR_ARG, "simulate: output buffers of repeat * Np samples needed (" +
                                            std::to_string(p.signal_size) + ")");
    const int lg = log2_of(p.fftpoint);
    if (!sim_plan_supported(lg))
        return set_error(MSOSM_ERR_UNSUPPORTED, "simulate: FFT length 2^" + std::to_string(lg) +
                                                    " (supported: 2^12..2^22 and 2^24..2^44)");

    // host: noise periods and chirp, exactly as the reference builds them
    std::vector<float> periods(2 * (size_t)p.periods_len), H(2 * (size_t)p.fftpoint);
    sim_noise(p, periods.data());
    sim_chirp(bw, dm, f0, p.fftpoint, H.data(), host_threads());

    if (device < 0) SIMCHK(hipGetDevice(&device)); // the caller's current device
    SIMCHK(hipSetDevice(device));
    SimDev d;
    SIMCHK(hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking));
    int cus = 0;
    SIMCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const long long N = (long long)p.fftpoint;
    float2 *dper = nullptr, *dH = nullptr, *dX = nullptr, *dc64 = nullptr;
    uint32_t *du16 = nullptr;
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dper), p.periods_len * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dH), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dX), N * sizeof(float2)));
    if (out_c64) SIMCHK(d.alloc(reinterpret_cast<void **>(&dc64), p.signal_size * sizeof(float2)));
    if (out_u16) SIMCHK(d.alloc(reinterpret_cast<void **>(&du16), p.signal_size * sizeof(uint32_t)));
    SimPlan plan;
    SIMCHK(make_plan(d, lg, cus, &plan));
    SIMCHK(hipMemcpyAsync(dper, periods.data(), p.periods_len * sizeof(float2), hipMemcpyHostToDevice, d.st));
    SIMCHK(hipMemcpyAsync(dH, H.data(), N * sizeof(float2), hipMemcpyHostToDevice, d.st));

    const long long per_chunk = (long long)SIM_PERIODS_PER_CHUNK * p.Np;
    for (unsigned long r = 0; r < p.mul_repeat; ++r) {
        // forward transform times the chirp, then the inverse with the extraction
        const long long done = (long long)r * per_chunk;
        SimFftArgs a{};
        a.chirp = dH;
        SIMCHK(run_plan(plan, dper + done, dX, -1, SIM_CHIRP, a, d.st));
        a.ex_start = (long long)p.start;
        a.ex_count = std::min<long long>(per_chunk, (long long)p.signal_size - done);
        a.ex_scale = 1.0f / static_cast<float>(p.fftpoint);
        a.ex_c64 = dc64 ? dc64 + done : nullptr;
        a.ex_u16 = du16 ? du16 + done : nullptr;
        SIMCHK(run_plan(plan, dX, nullptr, +1, SIM_EXTRACT, a, d.st));
    }
    if (out_c64) SIMCHK(hipMemcpyAsync(out_c64, dc64, p.signal_size * sizeof(float2), hipMemcpyDeviceToHost, d.st));
    if (out_u16) SIMCHK(hipMemcpyAsync(out_u16, du16, p.signal_size * sizeof(uint32_t), hipMemcpyDeviceToHost, d.st));
    SIMCHK(hipStreamSynchronize(d.st));
    return MSOSM_OK;
}

// ---- generate_pulsar_signal_new (simulated_complex_signal.cpp:279-467): the whole record of
// repeat + 2 guard periods dispersed by one FFT through a causal impulse response

extern "C" int msosm_simulate_params_new(float bw, float dm, float f0, float period, unsigned long repeat,
                                         unsigned long long *out)
{
    SimParamsNew p;
    if (!out) return set_error(MSOSM_ERR_ARG, "simulate_params_new: null output");
    if (sim_params_new(bw, dm, f0, period, repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_params_new: bw, f0, period > 0, dm >= 0 and at least one period");
    const unsigned long long v[MSOSM_SIM_NPARAMS] = {p.Np,          p.Nd,           p.range, p.fftpoint, p.signal_size,
                                                     p.filter_point, p.total_periods, p.start};
    std::copy(v, v + MSOSM_SIM_NPARAMS, out);
    return MSOSM_OK;
}

extern "C" int msosm_simulate_noise_new(float bw, float dm, float f0, float period, unsigned long repeat,
                                        float *record, unsigned long long len)
{
    SimParamsNew p;
    if (sim_params_new(bw, dm, f0, period, repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_noise_new: bw, f0, period > 0, dm >= 0 and at least one period");
    if (!record || len < p.fftpoint) return set_error(MSOSM_ERR_ARG, "simulate_noise_new: buffer too small");
    sim_noise_new(p, record);
    return MSOSM_OK;
}

extern "C" int msosm_simulate_pulsar_new(float bw, float dm, float f0, float period, unsigned long repeat, int device,
                                         uint32_t *out_u16, float *out_c64, unsigned long long len)
{
    SimParamsNew p;
    if (sim_params_new(bw, dm, f0, period, repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_new: bw, f0, period > 0, dm >= 0 and at least one period");
    if ((!out_u16 && !out_c64) || len < p.signal_size)
        return set_error(MSOSM_ERR_ARG, "simulate_new: output buffers of repeat * Np samples needed (" +
                                            std::to_string(p.signal_size) + ")");
    const int lg = log2_of(p.fftpoint), lgf = log2_of(p.filter_point);
    if (!sim_plan_supported(lg))
        return set_error(MSOSM_ERR_UNSUPPORTED, "simulate_new: FFT length 2^" + std::to_string(lg) +
                                                    " (supported: 2^12..2^22 and 2^24..2^44)");
    // the filter design's inverse FFT runs on the device where its length is supported; shorter
    // filters and 2^23 points run in FP64 on the host (filter_point <= fftpoint, so nothing else)
    const bool dev_filter = sim_plan_supported(lgf);
    const long long N = (long long)p.fftpoint, F = (long long)p.filter_point, Nd = (long long)p.Nd;

    // host: the chirp on the filter grid (:338-366) and the seed-1 record (:330-336)
    std::vector<float> Hf(2 * (size_t)F), record(2 * (size_t)N);
    sim_chirp(bw, dm, f0, p.filter_point, Hf.data(), host_threads());
    sim_noise_new(p, record.data());
    std::vector<float2> h_host;
    if (!dev_filter) {
        // :368-379: h = IFFT(H_filter) / filter_point
        std::vector<std::complex<double>> x(F);
        for (long long i = 0; i < F; ++i) x[i] = {Hf[2 * i], Hf[2 * i + 1]};
        const Fft64Handle *fft = design_fft_create((int)F);
        design_fft_run(fft, x.data(), +1);
        design_fft_destroy(fft);
        const float inv = 1.0f / static_cast<float>(F);
        h_host.resize(F);
        for (long long i = 0; i < F; ++i)
            h_host[i] = make_float2((float)x[i].real() * inv, (float)x[i].imag() * inv);
    }

    if (device < 0) SIMCHK(hipGetDevice(&device)); // the caller's current device
    SIMCHK(hipSetDevice(device));
    SimDev d;
    SIMCHK(hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking));
    int cus = 0;
    SIMCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    float2 *drec = nullptr, *dtap = nullptr, *dH = nullptr, *dh = nullptr, *dc64 = nullptr;
    uint32_t *du16 = nullptr;
    SIMCHK(d.alloc(reinterpret_cast<void **>(&drec), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dtap), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dH), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dh), F * sizeof(float2)));
    if (out_c64) SIMCHK(d.alloc(reinterpret_cast<void **>(&dc64), p.signal_size * sizeof(float2)));
    if (out_u16) SIMCHK(d.alloc(reinterpret_cast<void **>(&du16), p.signal_size * sizeof(uint32_t)));
    SimPlan plan;
    SIMCHK(make_plan(d, lg, cus, &plan));
    SIMCHK(hipMemcpyAsync(drec, record.data(), N * sizeof(float2), hipMemcpyHostToDevice, d.st));

    if (dev_filter) {
        // h = IFFT(H_filter) times 1/filter_point, through the extraction stage into dh
        SimPlan fplan;
        SIMCHK(make_plan(d, lgf, cus, &fplan));
        float2 *dHf = nullptr;
        SIMCHK(d.alloc(reinterpret_cast<void **>(&dHf), F * sizeof(float2)));
        SIMCHK(hipMemcpyAsync(dHf, Hf.data(), F * sizeof(float2), hipMemcpyHostToDevice, d.st));
        SimFftArgs a{};
        a.ex_start = 0;
        a.ex_count = F;
        a.ex_scale = 1.0f / static_cast<float>(F);
        a.ex_c64 = dh;
        SIMCHK(run_plan(fplan, dHf, nullptr, +1, SIM_EXTRACT, a, d.st));
    } else {
        SIMCHK(hipMemcpyAsync(dh, h_host.data(), F * sizeof(float2), hipMemcpyHostToDevice, d.st));
    }
    // :381-396: the causal response, the last Nd taps then h[0] at index Nd, zero-padded to fftpoint
    // (filter_point >= Nd + 1 and fftpoint >= Nd + 1, so both copies stay inside their buffers)
    SIMCHK(hipMemsetAsync(dtap, 0, N * sizeof(float2), d.st));
    if (Nd > 0) SIMCHK(hipMemcpyAsync(dtap, dh + (F - Nd), Nd * sizeof(float2), hipMemcpyDeviceToDevice, d.st));
    SIMCHK(hipMemcpyAsync(dtap + Nd, dh, sizeof(float2), hipMemcpyDeviceToDevice, d.st));
    SimFftArgs a{};
    SIMCHK(run_plan(plan, dtap, dH, -1, SIM_PLAIN, a, d.st)); // :398-400
    // :409-427: the record's spectrum times the response's, back, times 1/fftpoint; :431-455: the
    // repeat periods from guard * Np + Nd, and their uint16 pairs
    a.chirp = dH;
    SIMCHK(run_plan(plan, drec, dtap, -1, SIM_CHIRP, a, d.st));
    a.ex_start = (long long)p.start;
    a.ex_count = (long long)p.signal_size;
    a.ex_scale = 1.0f / static_cast<float>(p.fftpoint);
    a.ex_c64 = dc64;
    a.ex_u16 = du16;
    SIMCHK(run_plan(plan, dtap, nullptr, +1, SIM_EXTRACT, a, d.st));
    if (out_c64) SIMCHK(hipMemcpyAsync(out_c64, dc64, p.signal_size * sizeof(float2), hipMemcpyDeviceToHost, d.st));
    if (out_u16) SIMCHK(hipMemcpyAsync(out_u16, du16, p.signal_size * sizeof(uint32_t), hipMemcpyDeviceToHost, d.st));
    SIMCHK(hipStreamSynchronize(d.st));
    return MSOSM_OK;
}

// ---- generate_pulsar_signal (simulated_complex_signal.cpp:469-706): impulses of 32767 at Np/2 and
// Np/2 + 2 in each of range consecutive periods, dispersed through the response placed at
// [order, 2 order) of a 2 order FFT, the first Np samples summed over the periods and repeated

extern "C" int msosm_simulate_params_pulse(float bw, float dm, float f0, float period, unsigned long repeat,
                                           unsigned long long *out)
{
    SimParamsPulse p;
    if (!out) return set_error(MSOSM_ERR_ARG, "simulate_params_pulse: null output");
    if (sim_params_pulse(bw, dm, f0, period, repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_params_pulse: bw, f0, period > 0, dm > 0 and at least one period");
    const unsigned long long v[MSOSM_SIM_NPARAMS] = {p.Np, p.Nd, p.range, p.order, p.fftpoint, p.signal_size, 0, 0};
    std::copy(v, v + MSOSM_SIM_NPARAMS, out);
    return MSOSM_OK;
}

extern "C" int msosm_simulate_pulsar(float bw, float dm, float f0, float period, unsigned long repeat, int add_noise,
                                     float snr, int device, uint32_t *out_u16, float *out_c64, unsigned long long len)
{
    (void)snr;
    SimParamsPulse p;
    if (sim_params_pulse(bw, dm, f0, period, repeat, &p) != 0)
        return set_error(MSOSM_ERR_ARG, "simulate_pulsar: bw, f0, period > 0, dm > 0 and at least one period");
    if (add_noise)
        return set_error(MSOSM_ERR_UNSUPPORTED, "simulate_pulsar: add_noise (the reference's noise branch rewrites "
                                                "period 0 while other threads copy it, so its record is not defined)");
    if ((!out_u16 && !out_c64) || len < p.signal_size)
        return set_error(MSOSM_ERR_ARG, "simulate_pulsar: output buffers of repeat * Np samples needed (" +
                                            std::to_string(p.signal_size) + ")");
    const int lg = log2_of(p.fftpoint), lgo = log2_of(p.order);
    if (!sim_plan_supported(lg))
        return set_error(MSOSM_ERR_UNSUPPORTED, "simulate_pulsar: FFT length 2^" + std::to_string(lg) +
                                                    " (supported: 2^12..2^22 and 2^24..2^44)");
    const bool dev_filter = sim_plan_supported(lgo); // else FP64 on the host (order < 2^12 or 2^23)
    const long long N = (long long)p.fftpoint, O = (long long)p.order, Np = (long long)p.Np;

    // host: the chirp (:524-546), the impulse train (:596-600), and the response if designed here
    std::vector<float> H1(2 * (size_t)O);
    sim_chirp_pulse(bw, dm, f0, p.order, H1.data(), host_threads());
    std::vector<float2> train(N, make_float2(0.f, 0.f));
    for (unsigned long i = 0; i < p.range; ++i) {
        train[i * Np + Np / 2].x = 32767.0f;
        train[i * Np + Np / 2 + 2].x = 32767.0f;
    }
    std::vector<float2> h_host;
    if (!dev_filter) {
        // :548-557: IFFT, then each component divided by order
        std::vector<std::complex<double>> x(O);
        for (long long i = 0; i < O; ++i) x[i] = {H1[2 * i], H1[2 * i + 1]};
        const Fft64Handle *fft = design_fft_create((int)O);
        design_fft_run(fft, x.data(), +1);
        design_fft_destroy(fft);
        h_host.resize(O);
        for (long long i = 0; i < O; ++i)
            h_host[i] = make_float2((float)x[i].real() / (float)O, (float)x[i].imag() / (float)O);
    }

    // host buffers the device copies into: declared before d, whose destructor synchronises the
    // stream, so a copy still queued at an early return never lands in freed memory
    std::vector<float2> per(Np);
    std::vector<uint32_t> codes(Np);
    if (device < 0) SIMCHK(hipGetDevice(&device)); // the caller's current device
    SIMCHK(hipSetDevice(device));
    SimDev d;
    SIMCHK(hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking));
    int cus = 0;
    SIMCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    float2 *dtrain = nullptr, *dresp = nullptr, *dH = nullptr, *dper = nullptr;
    uint32_t *du16 = nullptr;
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dtrain), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dresp), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dH), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dper), Np * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&du16), Np * sizeof(uint32_t)));
    SimPlan plan;
    SIMCHK(make_plan(d, lg, cus, &plan));
    SIMCHK(hipMemcpyAsync(dtrain, train.data(), N * sizeof(float2), hipMemcpyHostToDevice, d.st));
    // :559-561: the response at [order, 2 order) of fftpoint zeros
    SIMCHK(hipMemsetAsync(dresp, 0, O * sizeof(float2), d.st));
    if (dev_filter) {
        SimPlan oplan;
        SIMCHK(make_plan(d, lgo, cus, &oplan));
        float2 *dH1 = nullptr;
        SIMCHK(d.alloc(reinterpret_cast<void **>(&dH1), O * sizeof(float2)));
        SIMCHK(hipMemcpyAsync(dH1, H1.data(), O * sizeof(float2), hipMemcpyHostToDevice, d.st));
        SimFftArgs a{};
        a.ex_start = 0;
        a.ex_count = O;
        a.ex_scale = 1.0f / static_cast<float>(O); // a power of two: the same as dividing by order
        a.ex_c64 = dresp + O;
        SIMCHK(run_plan(oplan, dH1, nullptr, +1, SIM_EXTRACT, a, d.st));
    } else {
        SIMCHK(hipMemcpyAsync(dresp + O, h_host.data(), O * sizeof(float2), hipMemcpyHostToDevice, d.st));
    }
    SimFftArgs a{};
    SIMCHK(run_plan(plan, dresp, dH, -1, SIM_PLAIN, a, d.st)); // :563-565
    // :588-611: the train's spectrum times the response's, times 1/fftpoint, back; the first Np
    // samples. The reference transforms each period's impulses separately and sums the results; the
    // transforms are linear, so one transform of the whole train gives the same sum.
    a.chirp = dH;
    SIMCHK(run_plan(plan, dtrain, dresp, -1, SIM_CHIRP, a, d.st));
    a.ex_start = 0;
    a.ex_count = Np;
    a.ex_scale = 1.0f / static_cast<float>(p.fftpoint);
    a.ex_c64 = dper;
    a.ex_u16 = du16;
    SIMCHK(run_plan(plan, dresp, nullptr, +1, SIM_EXTRACT, a, d.st));
    SIMCHK(hipMemcpyAsync(per.data(), dper, Np * sizeof(float2), hipMemcpyDeviceToHost, d.st));
    SIMCHK(hipMemcpyAsync(codes.data(), du16, Np * sizeof(uint32_t), hipMemcpyDeviceToHost, d.st));
    SIMCHK(hipStreamSynchronize(d.st));
    // :656-677: the period repeated
    for (unsigned long r = 0; r < p.repeat; ++r) {
        if (out_c64) std::copy(per.begin(), per.end(), reinterpret_cast<float2 *>(out_c64) + r * Np);
        if (out_u16) std::copy(codes.begin(), codes.end(), out_u16 + r * Np);
    }
    return MSOSM_OK;
}
