// Host side of the GPU pulsar simulator (sim_host.cpp).
#pragma once

namespace msosm {

constexpr unsigned long SIM_PERIODS_PER_CHUNK = 128; // periods per dispersion FFT (mul_count)

struct SimParams {
    unsigned long Np, Nd, range, guard; // period and dispersion samples, guard periods = range + 2
    unsigned long repeat, signal_size;  // periods simulated and samples returned (repeat * Np)
    unsigned long mul_repeat;           // chunks of 128 periods, one FFT each
    unsigned long fftpoint, offset;     // FFT length, first sample of the generated periods
    unsigned long dummy0, valid, dummy1; // periods from seed 1, from seed 2, after them
    unsigned long periods_len;          // complex samples of the noise buffer
    unsigned long start;                // first sample of the extracted part of a chunk's window
};

// generate_pulsar_signal_new: one continuous record, dispersed by one FFT through a causal filter
struct SimParamsNew {
    unsigned long Np, Nd, range, guard; // as above
    unsigned long repeat, signal_size;  // periods returned and their samples (repeat * Np)
    unsigned long total_periods;        // repeat + 2 guard periods generated (all from seed 1)
    unsigned long fftpoint;             // record FFT length >= total_periods * Np + Nd
    unsigned long filter_point;         // filter design length >= Nd + 1
    unsigned long start;                // first extracted sample: guard * Np + Nd
};

// generate_pulsar_signal: one period of a dispersed two-impulse pulse, repeated
struct SimParamsPulse {
    unsigned long Np, Nd, range; // range = ceil(Nd / Np) in float, as there
    unsigned long order;         // impulse-response length: next power of two >= range * Np
    unsigned long fftpoint;      // 2 order
    unsigned long repeat, signal_size;
};

int sim_params(float bw, float dm, float f0, float period, unsigned long dummy_repeat, unsigned long valid_repeat,
               SimParams *p);
int sim_params_pulse(float bw, float dm, float f0, float period, unsigned long repeat, SimParamsPulse *p);
// the chirp of generate_pulsar_signal on `order` bins: w = 2 pi fs i / order for i in [0, order), stored
// at (i + order / 2) mod order (:528-546)
void sim_chirp_pulse(float bw, float dm, float f0, unsigned long order, float *H, int nthreads);
void sim_noise(const SimParams &p, float *periods);
int sim_params_new(float bw, float dm, float f0, float period, unsigned long repeat, SimParamsNew *p);
// the record: total_periods periods from one seed-1 stream, zero-padded to fftpoint samples
void sim_noise_new(const SimParamsNew &p, float *record);
void sim_chirp(float bw, float dm, float f0, unsigned long fftpoint, float *H, int nthreads);

} // namespace msosm
