// Reference-compatible pulsar simulator class (simulated_pulsar/include/simulated_complex_signal.h:8-55)
// over the C ABI: generate_pulsar_signal_block disperses on the GPU (msosm_simulate_pulsar_block).
#pragma once

#include <string>

#include "globals.h"

class SimulatedComplexSignal {
public:
    SimulatedComplexSignal(float bw, float dm, float f0, float period, string mode = "complex");
    // simulated_complex_signal.cpp:74-277; the record is dummy_repeat + valid_repeat periods
    void generate_pulsar_signal_block(unsigned long dummy_repeat = 0, unsigned long valid_repeat = 1);
    ~SimulatedComplexSignal();

public:
    // the simulated signal: only the uint16 form is produced (as by the block generator)
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
    float bw, fs, kdm, dm, f0, w0, period, signal_power = 0.f;
};
