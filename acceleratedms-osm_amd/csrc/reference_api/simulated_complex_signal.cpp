// Reference-compatible SimulatedComplexSignal (simulated_pulsar/simulated_complex_signal.cpp:16-41,
// 74-277, 279-467, 469-706): the constructor sets the reference's members; the three generators return
// the same records (noise and chirp as the reference builds them,
// dispersion on the current GPU).
#include "simulated_complex_signal.h"

#include <cstdlib>
#include <cstring>

SimulatedComplexSignal::SimulatedComplexSignal(float bw, float dm, float f0, float period, string mode)
    : data_type(mode), bw(bw), fs(bw), kdm((float)(4.15 * 1e15)), dm(dm), f0(f0), w0((float)(2 * pi * f0)),
      period(period)
{
    unsigned long long p[MSOSM_SIM_NPARAMS];
    msosm::check(msosm_simulate_params(bw, dm, f0, period, 0, 1, p), "SimulatedComplexSignal");
    Np = (unsigned long)p[0];
    Nd = (unsigned long)p[1];
    if (verbose) {
        cout << "Test signal parameters:" << endl;
        cout << "Bandwidth: " << bw << "Hz" << endl;
        cout << "Dispersion measure: " << dm << "pc cm^-3" << endl;
        cout << "Start frequency: " << f0 << "Hz" << endl;
        cout << "Pulsar period: " << period << "s" << endl;
        cout << "Period samples: " << Np << endl;
        cout << "Dispersion samples: " << Nd << endl;
        line();
    }
}

void SimulatedComplexSignal::generate_pulsar_signal_block(unsigned long dummy_repeat, unsigned long valid_repeat)
{
    if (data_type != "uint16") {
        cout << "Warning: this function is intended for uint16 data type only." << endl;
        return;
    }
    unsigned long long p[MSOSM_SIM_NPARAMS];
    msosm::check(msosm_simulate_params(bw, dm, f0, period, dummy_repeat, valid_repeat, p),
                 "generate_pulsar_signal_block");
    range = (unsigned long)p[2];
    signal_size = (unsigned long)p[4];
    // whole chunks of 128 periods, as the reference allocates; the record is the first signal_size
    const size_t alloc = (size_t)p[7] * 128 * Np;
    release();
    signal_u16 = (uint16_pair *)malloc(sizeof(uint16_pair) * alloc);
    if (!signal_u16) msosm::check(MSOSM_ERR_ALLOC, "generate_pulsar_signal_block: host buffer");
    memset(signal_u16 + signal_size, 0, sizeof(uint16_pair) * (alloc - signal_size));
    msosm::check(msosm_simulate_pulsar_block(bw, dm, f0, period, dummy_repeat, valid_repeat, -1,
                                             reinterpret_cast<uint32_t *>(signal_u16), nullptr, signal_size),
                 "generate_pulsar_signal_block");
}

void SimulatedComplexSignal::generate_pulsar_signal_new(unsigned long repeat)
{
    unsigned long long p[MSOSM_SIM_NPARAMS];
    msosm::check(msosm_simulate_params_new(bw, dm, f0, period, repeat, p), "generate_pulsar_signal_new");
    range = (unsigned long)p[2];
    signal_size = (unsigned long)p[4];
    release();
    signal = (fftwf_complex *)malloc(sizeof(fftwf_complex) * signal_size);
    if (data_type == "uint16") signal_u16 = (uint16_pair *)malloc(sizeof(uint16_pair) * signal_size);
    if (!signal || (data_type == "uint16" && !signal_u16))
        msosm::check(MSOSM_ERR_ALLOC, "generate_pulsar_signal_new: host buffer");
    msosm::check(msosm_simulate_pulsar_new(bw, dm, f0, period, repeat, -1, reinterpret_cast<uint32_t *>(signal_u16),
                                           reinterpret_cast<float *>(signal), signal_size),
                 "generate_pulsar_signal_new");
}

void SimulatedComplexSignal::generate_pulsar_signal(unsigned long repeat, bool add_noise, float SNR, bool pinned)
{
    unsigned long long p[MSOSM_SIM_NPARAMS];
    msosm::check(msosm_simulate_params_pulse(bw, dm, f0, period, repeat, p), "generate_pulsar_signal");
    range = (unsigned long)p[2];
    signal_size = (unsigned long)p[5];
    release();
    const bool u16 = data_type == "uint16";
    if (pinned) {
        pinned_ = true;
        msosm::check(msosm_host_alloc(sizeof(fftwf_complex) * signal_size, reinterpret_cast<void **>(&signal)),
                     "generate_pulsar_signal");
        if (u16)
            msosm::check(msosm_host_alloc(sizeof(uint16_pair) * signal_size, reinterpret_cast<void **>(&signal_u16)),
                         "generate_pulsar_signal");
    } else {
        signal = (fftwf_complex *)malloc(sizeof(fftwf_complex) * signal_size);
        if (u16) signal_u16 = (uint16_pair *)malloc(sizeof(uint16_pair) * signal_size);
        if (!signal || (u16 && !signal_u16)) msosm::check(MSOSM_ERR_ALLOC, "generate_pulsar_signal: host buffer");
    }
    msosm::check(msosm_simulate_pulsar(bw, dm, f0, period, repeat, add_noise ? 1 : 0, SNR, -1,
                                       reinterpret_cast<uint32_t *>(signal_u16), reinterpret_cast<float *>(signal),
                                       signal_size),
                 "generate_pulsar_signal");
}

void SimulatedComplexSignal::release()
{
    if (pinned_) {
        (void)msosm_host_free(signal);
        (void)msosm_host_free(signal_u16);
    } else {
        free(signal);
        free(signal_u16);
    }
    signal = nullptr;
    signal_u16 = nullptr;
    pinned_ = false;
}

SimulatedComplexSignal::~SimulatedComplexSignal() { release(); }
