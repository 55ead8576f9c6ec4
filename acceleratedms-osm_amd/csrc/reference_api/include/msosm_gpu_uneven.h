// MSOSM_GPU_UNEVEN: the MS-OSM device pipeline for DM grids whose trials keep their own FFT length
// (uneven-M construction; the reference gives every DM the grid's largest M, msosm_prepare.cpp:67).
// Same constructor and call pattern as MSOSM_GPU_BATCH (dedispersion/include/msosm_gpu_batch.h:13-25)
// over the msosm_multi_* C ABI: all DMs dedisperse the same push of samples, each at its own length,
// by default in one persistent launch.
#pragma once

#include <vector>

#include "globals.h"

class Fold_GPU;

class MSOSM_GPU_UNEVEN {
public:
    MSOSM_GPU_UNEVEN(float bw, float *dm, float f0, int numDMs);
    ~MSOSM_GPU_UNEVEN();
    // push_blocks blocks of the largest M per push; MSOSM_IN_* input, MSOSM_OUT_* outputs
    void initialize(int push_blocks = 1, int in_format = MSOSM_IN_C64, int out_flags = MSOSM_OUT_C64,
                    int phase_mode = MSOSM_PHASE_FP64);
    void filter_block_uint16(uint16_pair *input);
    void filter_block_int8(const int8_t *iq);
    void filter_block_c64(const Complex *x);
    void filter_block_device(const void *dev_input);
    // complex64 series of DM i for the last push (push_samples() values)
    void get_output(int dm_index, Complex *output);
    Complex *get_output_pointer(int dm_index);
    // fold every DM to nbins at period (one profile per DM, read with get_folded_data)
    void fold(float period, unsigned long nbins);
    std::vector<float> get_folded_data();
    void set_launch_mode(int mode);
    void set_graph(bool on);
    void synchronize();

    long long push_samples() const { return push_; }
    int num_classes() const { return classes_; }

    int numDMs;
    std::vector<float> dm_values;
    std::vector<int> M_values, delaycount_values;

private:
    float bw_, f0_;
    long long push_ = 0;
    int classes_ = 0, in_format_ = MSOSM_IN_C64;
    unsigned long nbins_ = 0;
    msosm_multi *multi_ = nullptr;
    void push(const void *input);
};
