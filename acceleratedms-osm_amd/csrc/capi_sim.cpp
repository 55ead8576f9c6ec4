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

int host_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

} // namespace

#define SIMCHK(x)                                                                                                      \
    do {                                                                                                               \
        const hipError_t e_ = (x);                                                                                     \
        if (e_ != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("simulate: ") + hipGetErrorString(e_));     \
    } while (0)

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
    if (!H || fftpoint < 2 || (fftpoint & (fftpoint - 1))) return set_error(MSOSM_ERR_ARG, "simulate_chirp: bad length");
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
    int lg = 0;
    while ((1UL << lg) < p.fftpoint) ++lg;
    // one workgroup per transform up to 2^22; the six-step form over two supported lengths above
    const bool one = sim_length_supported(lg);
    const int lga = (lg + 1) / 2, lgb = lg / 2;
    if (!one && !(sim_length_supported(lga) && sim_length_supported(lgb)))
        return set_error(MSOSM_ERR_UNSUPPORTED, "simulate: FFT length 2^" + std::to_string(lg) +
                                                    " (supported: 2^12..2^22 and 2^24..2^44)");
    const long long N = (long long)p.fftpoint, A = 1LL << lga, B = 1LL << lgb;

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
    float2 *dper = nullptr, *dH = nullptr, *dX = nullptr, *dT = nullptr, *scr = nullptr, *dc64 = nullptr;
    uint32_t *du16 = nullptr;
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dper), p.periods_len * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dH), N * sizeof(float2)));
    SIMCHK(d.alloc(reinterpret_cast<void **>(&dX), N * sizeof(float2)));
    if (!one) SIMCHK(d.alloc(reinterpret_cast<void **>(&dT), N * sizeof(float2)));
    if (out_c64) SIMCHK(d.alloc(reinterpret_cast<void **>(&dc64), p.signal_size * sizeof(float2)));
    if (out_u16) SIMCHK(d.alloc(reinterpret_cast<void **>(&du16), p.signal_size * sizeof(uint32_t)));
    const long long Lsub = one ? N : A; // A >= B
    const int grid = one ? 1 : std::max(8, cus);
    SIMCHK(d.alloc(reinterpret_cast<void **>(&scr), (size_t)grid * Lsub * sizeof(float2)));
    SIMCHK(hipMemcpyAsync(dper, periods.data(), p.periods_len * sizeof(float2), hipMemcpyHostToDevice, d.st));
    SIMCHK(hipMemcpyAsync(dH, H.data(), N * sizeof(float2), hipMemcpyHostToDevice, d.st));
    SimTw ta, tb;
    SIMCHK(make_tw(d, one ? lg : lga, &ta));
    if (!one) SIMCHK(make_tw(d, lgb, &tb));

    const long long per_chunk = (long long)SIM_PERIODS_PER_CHUNK * p.Np;
    for (unsigned long r = 0; r < p.mul_repeat; ++r) {
        const float2 *window = dper + r * per_chunk;
        const long long done = (long long)r * per_chunk;
        SimFftArgs a{};
        a.scratch = scr;
        a.ex_start = (long long)p.start;
        a.ex_count = std::min<long long>(per_chunk, (long long)p.signal_size - done);
        a.ex_scale = 1.0f / static_cast<float>(p.fftpoint);
        a.ex_c64 = dc64 ? dc64 + done : nullptr;
        a.ex_u16 = du16 ? du16 + done : nullptr;
        a.chirp = dH;
        a.big_n = N;
        if (one) {
            // forward transform times the chirp, then the inverse with the extraction
            set_tw(a, ta);
            a.ntrans = 1;
            a.in_t = a.out_t = 0;
            a.in_e = a.out_e = 1;
            a.src = window;
            a.dst = dX;
            a.mode = SIM_CHIRP;
            SIMCHK(launch_sim_fft(lg, -1, a, grid, d.st));
            a.src = dX;
            a.mode = SIM_EXTRACT;
            SIMCHK(launch_sim_fft(lg, +1, a, grid, d.st));
        } else {
            // six-step: N = A B, n = c + B a, k = a' + A c'. Columns (length A, stride B) with the
            // twiddle W_N^{a' c}, then rows (length B) written transposed in natural order.
            auto cols = [&](const float2 *src, float2 *dst, int dir) {
                set_tw(a, ta);
                a.ntrans = B;
                a.in_t = a.out_t = 1;
                a.in_e = a.out_e = B;
                a.src = src;
                a.dst = dst;
                a.mode = SIM_TWIDDLE;
                return launch_sim_fft(lga, dir, a, grid, d.st);
            };
            auto rows = [&](const float2 *src, float2 *dst, int dir, int mode) {
                set_tw(a, tb);
                a.ntrans = A;
                a.in_t = B;
                a.in_e = 1;
                a.out_t = 1;
                a.out_e = A;
                a.src = src;
                a.dst = dst;
                a.mode = mode;
                return launch_sim_fft(lgb, dir, a, grid, d.st);
            };
            SIMCHK(cols(window, dT, -1));
            SIMCHK(rows(dT, dX, -1, SIM_CHIRP));
            SIMCHK(cols(dX, dT, +1));
            SIMCHK(rows(dT, nullptr, +1, SIM_EXTRACT));
        }
    }
    if (out_c64) SIMCHK(hipMemcpyAsync(out_c64, dc64, p.signal_size * sizeof(float2), hipMemcpyDeviceToHost, d.st));
    if (out_u16) SIMCHK(hipMemcpyAsync(out_u16, du16, p.signal_size * sizeof(uint32_t), hipMemcpyDeviceToHost, d.st));
    SIMCHK(hipStreamSynchronize(d.st));
    return MSOSM_OK;
}
