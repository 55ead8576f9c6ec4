// Host side of the GPU pulsar simulator: the parameters, the noise periods and the dispersion chirps of
// SimulatedComplexSignal's three generators (simulated_complex_signal.cpp:16-41, 43-72, 74-195,
// 279-345, 469-546), with the reference's types and expression order (compiled without FMA
// contraction). The
// random numbers come from the same standard-library engines (std::mt19937, std::normal_distribution),
// so the noise is the reference's bit for bit.
#include "sim_host.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <thread>
#include <vector>

namespace msosm {

int sim_params(float bw, float dm, float f0, float period, unsigned long dummy_repeat, unsigned long valid_repeat,
               SimParams *p)
{
    if (!(bw > 0.f) || !(f0 > 0.f) || !(period > 0.f) || !(dm >= 0.f) || valid_repeat + dummy_repeat == 0) return -1;
    // constructor (simulated_complex_signal.cpp:16-29): float members fs, kdm, f0, bw, dm, period
    const float fs = bw;
    const float kdm = (float)(4.15 * 1e15);
    p->Np = (unsigned long)(period * fs);
    if (p->Np == 0) return -1;
    const double inner = 1 / std::pow((double)f0, 2) - 1 / std::pow((double)(float)(f0 + bw), 2);
    const float inv_fs = 1 / fs;
    p->Nd = (unsigned long)std::floor((double)(float)(kdm * dm) * inner / (double)inv_fs);
    // generate_pulsar_signal_block (:74-131)
    p->repeat = dummy_repeat + valid_repeat;
    p->signal_size = p->repeat * p->Np;
    const unsigned long mul_count = SIM_PERIODS_PER_CHUNK;
    p->mul_repeat = (unsigned long)std::ceil((double)p->repeat / (double)mul_count);
    p->range = (unsigned long)std::ceil((double)p->Nd / (double)p->Np);
    p->guard = p->range + 2;
    const unsigned long num_periods = 2 * p->guard + mul_count;
    const unsigned long long_size = num_periods * p->Np;
    p->fftpoint = 1;
    while (p->fftpoint < long_size) p->fftpoint <<= 1;
    p->offset = p->fftpoint - long_size;
    p->dummy0 = p->guard + dummy_repeat - 1;
    p->valid = valid_repeat;
    p->dummy1 = (num_periods + p->repeat - mul_count) - p->dummy0 - p->valid;
    // every chunk's window of fftpoint samples lies inside the buffer; the reference sizes it
    // fftpoint + (repeat - 128) Np, which only covers the last window when repeat is a multiple of 128
    p->periods_len = p->fftpoint + (p->mul_repeat - 1) * mul_count * p->Np;
    p->start = p->offset + p->guard * p->Np;
    return 0;
}

// generate_test_pulse (:43-72): one period of amplitude-modulated complex noise
static void test_pulse(float *sig, const std::vector<double> &amp, std::mt19937 &gen)
{
    std::normal_distribution<double> normal_dist(0.0, 1.0); // a new distribution per period, as there
    for (size_t i = 0; i < amp.size(); ++i) {
        sig[2 * i] = static_cast<float>(amp[i] * normal_dist(gen));
        sig[2 * i + 1] = static_cast<float>(amp[i] * normal_dist(gen));
    }
}

// the three-Gaussian profile's per-component amplitude (:47-63), the same for every period
static std::vector<double> pulse_amp(unsigned long Np)
{
    std::vector<double> profile(Np), amp(Np);
    double max_amp = 0.0;
    for (unsigned long i = 0; i < Np; i++) {
        const double x = static_cast<double>(i) / static_cast<double>(Np);
        const double p1 = 0.45 * std::exp(-0.5 * std::pow((x - 0.455) / 0.008, 2.0));
        const double p2 = 1.00 * std::exp(-0.5 * std::pow((x - 0.500) / 0.012, 2.0));
        const double p3 = 0.28 * std::exp(-0.5 * std::pow((x - 0.555) / 0.006, 2.0));
        profile[i] = p1 + p2 + p3;
        max_amp = std::max(max_amp, profile[i]);
    }
    const double scale = 10000.0 / max_amp;
    for (unsigned long i = 0; i < Np; i++) amp[i] = profile[i] * scale / std::sqrt(2.0);
    return amp;
}

int sim_params_new(float bw, float dm, float f0, float period, unsigned long repeat, SimParamsNew *p)
{
    SimParams b; // the constructor's Np and Nd
    if (repeat == 0 || sim_params(bw, dm, f0, period, 0, repeat, &b) != 0) return -1;
    // generate_pulsar_signal_new (:279-328)
    p->Np = b.Np;
    p->Nd = b.Nd;
    p->repeat = repeat;
    p->signal_size = repeat * b.Np;
    p->range = (unsigned long)std::ceil((double)b.Nd / (double)b.Np);
    p->guard = p->range + 2;
    p->total_periods = repeat + 2 * p->guard;
    const unsigned long required = p->total_periods * b.Np + b.Nd;
    p->fftpoint = 1;
    while (p->fftpoint < required) p->fftpoint <<= 1;
    // :343-345
    p->filter_point = 1;
    while (p->filter_point < b.Nd + 1) p->filter_point <<= 1;
    p->start = p->guard * b.Np + b.Nd; // :431
    return 0;
}

void sim_noise_new(const SimParamsNew &p, float *record)
{
    // :330-336: every period from one seed-1 stream, the rest of the FFT window zero
    std::fill(record + 2 * p.total_periods * p.Np, record + 2 * p.fftpoint, 0.f);
    const std::vector<double> amp = pulse_amp(p.Np);
    std::mt19937 gen(1);
    for (unsigned long q = 0; q < p.total_periods; q++) test_pulse(record + 2 * q * p.Np, amp, gen);
}

void sim_noise(const SimParams &p, float *periods)
{
    std::fill(periods, periods + 2 * p.periods_len, 0.f);
    const std::vector<double> amp = pulse_amp(p.Np);
    // periods: dummy0 from seed 1, then valid and dummy1 from seed 2 (:119-161, no cached file)
    std::mt19937 gen(1);
    unsigned long q = 0;
    for (unsigned long k = 0; k < p.dummy0; k++, q++) test_pulse(periods + 2 * (p.offset + q * p.Np), amp, gen);
    gen.seed(2);
    for (unsigned long k = 0; k < p.valid + p.dummy1; k++, q++)
        test_pulse(periods + 2 * (p.offset + q * p.Np), amp, gen);
}

int sim_params_pulse(float bw, float dm, float f0, float period, unsigned long repeat, SimParamsPulse *p)
{
    SimParams b; // the constructor's Np and Nd
    if (repeat == 0 || sim_params(bw, dm, f0, period, 0, repeat, &b) != 0) return -1;
    p->Np = b.Np;
    p->Nd = b.Nd;
    p->repeat = repeat;
    p->signal_size = repeat * b.Np;
    p->range = (unsigned long)std::ceil((float)b.Nd / (float)b.Np); // :500, in float
    if (p->range == 0) return -1; // no dispersion: the reference's order would be 2^-inf
    if (b.Np < 5) return -1; // the second impulse, at Np / 2 + 2, lies inside the period
    // :506-507: 2^ceil(log2(range Np)), the next power of two
    p->order = 1;
    while (p->order < p->range * p->Np) p->order <<= 1;
    p->fftpoint = 2 * p->order;
    return 0;
}

void sim_chirp_pulse(float bw, float dm, float f0, unsigned long order, float *H, int nthreads)
{
    const double pi_d = 3.14159265358979323846;
    const double kdm_d = 4.15e15;
    const double fs_d = static_cast<double>(bw);
    const double dm_d = static_cast<double>(dm);
    const double f0_d = static_cast<double>(f0);
    const double w0_d = 2.0 * pi_d * f0_d;
    const double step = 2.0 * pi_d * fs_d / static_cast<double>(order);
    auto work = [&](unsigned long lo, unsigned long hi) {
        for (unsigned long i = lo; i < hi; i++) {
            const double w = step * static_cast<double>(i);
            const double phase = 4.0 * pi_d * pi_d * kdm_d * dm_d * w * w / ((w + w0_d) * w0_d * w0_d);
            const unsigned long k = i < order / 2 ? order / 2 + i : i - order / 2;
            H[2 * k] = static_cast<float>(std::cos(phase));
            H[2 * k + 1] = static_cast<float>(std::sin(phase));
        }
    };
    const int n = std::max(1, std::min(nthreads, 64));
    std::vector<std::thread> th;
    const unsigned long chunk = (order + n - 1) / n;
    for (int i = 0; i < n; ++i) {
        const unsigned long lo = std::min<unsigned long>(order, i * chunk), hi = std::min<unsigned long>(order, lo + chunk);
        if (lo < hi) th.emplace_back(work, lo, hi);
    }
    for (auto &t : th) t.join();
}

void sim_chirp(float bw, float dm, float f0, unsigned long fftpoint, float *H, int nthreads)
{
    // :166-198, FP64 phase rounded to complex64 (fs = bw as a float member)
    const double pi_d = 3.14159265358979323846;
    const double kdm_d = 4.15e15;
    const double fs_d = static_cast<double>(bw);
    const double dm_d = static_cast<double>(dm);
    const double f0_d = static_cast<double>(f0);
    const double w0_d = 2.0 * pi_d * f0_d;
    const double freq_step = fs_d / static_cast<double>(fftpoint);
    auto work = [&](unsigned long lo, unsigned long hi) {
        for (unsigned long k = lo; k < hi; k++) {
            double f_disp;
            if (k < fftpoint / 2)
                f_disp = fs_d / 2.0 + static_cast<double>(k) * freq_step;
            else
                f_disp = static_cast<double>(k) * freq_step - fs_d / 2.0;
            const double w = 2.0 * pi_d * f_disp;
            const double phase = 4.0 * pi_d * pi_d * kdm_d * dm_d * w * w / ((w + w0_d) * w0_d * w0_d);
            H[2 * k] = static_cast<float>(std::cos(phase));
            H[2 * k + 1] = static_cast<float>(std::sin(phase));
        }
    };
    const int n = std::max(1, std::min(nthreads, 64));
    std::vector<std::thread> th;
    const unsigned long step = (fftpoint + n - 1) / n;
    for (int i = 0; i < n; ++i) {
        const unsigned long lo = std::min<unsigned long>(fftpoint, i * step), hi = std::min<unsigned long>(fftpoint, lo + step);
        if (lo < hi) th.emplace_back(work, lo, hi);
    }
    for (auto &t : th) t.join();
}

} // namespace msosm
