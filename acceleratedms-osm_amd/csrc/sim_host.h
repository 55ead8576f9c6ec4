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

int sim_params(float bw, float dm, float f0, float period, unsigned long dummy_repeat, unsigned long valid_repeat,
               SimParams *p);
void sim_noise(const SimParams &p, float *periods);
void sim_chirp(float bw, float dm, float f0, unsigned long fftpoint, float *H, int nthreads);

} // namespace msosm
