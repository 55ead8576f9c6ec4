// Reference-compatible pulsar simulator class (simulated_pulsar/include/simulated_complex_signal.h:8-55)
// over the C ABI: the three generators disperse on the GPU (msosm_simulate_pulsar, _block, _new).
#pragma once

#include <string>

#include "globals.h"

class SimulatedComplexSignal {
public:
    SimulatedComplexSignal(float bw, float dm, float f0, float period, string mode = "complex");
    // simulated_complex_signal.cpp:74-277; the record is dummy_repeat + valid_repeat periods
    void generate_pulsar_signal_block(unsigned long dummy_repeat = 0, unsigned long valid_repeat = 1);
    // simulated_complex_signal.cpp:279-467; fills signal, and signal_u16 in "uint16" mode
    void generate_pulsar_signal_new(unsigned long repeat = 1);
    // simulated_complex_signal.cpp:469-706; add_noise must be false (as msosm_save_simulated.cpp:115 and
    // test_msosm_speed.cpp:160 pass), pinned selects page-locked host buffers
    void generate_pulsar_signal(unsigned long repeat = 1, bool add_noise = false, float SNR = 0, bool pinned = true);
    ~SimulatedComplexSignal();

public:
    // the simulated signal: the block generator fills signal_u16 only, as the reference's does
    fftwf_complex *signal = nullptr;
    uint16_pair *signal_u16 = nullptr;
    string data_type;
    unsigned long signal_size = 0;
    unsigned long Np = 0;
    unsigned long Nd = 0;
    unsigned long M = 0;
    unsigned long M_min = 0;
    unsigned long range = 0;

private:
    void release();
    bool pinned_ = false; // signal and signal_u16 from msosm_host_alloc (page-locked)
    float bw, fs, kdm, dm, f0, w0, period, signal_power = 0.f;
};
