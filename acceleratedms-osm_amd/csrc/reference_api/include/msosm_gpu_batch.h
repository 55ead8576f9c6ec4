// MSOSM_GPU_BATCH: the device pipeline of the MS-OSM method
// (dedispersion/include/msosm_gpu_batch.h:13-97, dedispersion/msosm/msosm_gpu_batch.cpp), run on the
// MI355X HIP kernels through the C ABI of msosm_amd.h.
#pragma once

#include <vector>

#include "msosm_prepare.h"

// Segment runs of equal delay (gpu/include/gpu_kernel.cuh:9-21); the HIP kernels read a per-bin delay
// table instead, but the runs stay available for callers that inspect the design.
typedef struct {
    int start;
    int len;
    int group;
} Segment;

constexpr int BLOCK_SIZE = 256;
constexpr int ITEMS_PER_THREAD = 4;

class Fold_GPU;

class MSOSM_GPU_BATCH : public Prepare_MSOSM {
public:
    MSOSM_GPU_BATCH(float bw = 0, float dm = 0, float f0 = 0);
    MSOSM_GPU_BATCH(float bw, float *dm, float f0, int numDMs = 1);
    ~MSOSM_GPU_BATCH() override;
    void get_device_info();
    // fold == true: complex64 output for get_output_pointer(); fold == false: uint16 output for get_output()
    void initialize_uint16(int fftpoint = 0, int batch = 1, bool fold = false);
    void filter_block_uint16(uint16_pair *input);
    void get_output(uint16_pair *output);
    void synchronize();
    // releases this object's device memory (the reference reset the whole device)
    void reset_device();
    Complex *get_output_pointer();

    // ---- extensions over the reference interface ----
    // any input format (MSOSM_IN_*) and any output set (MSOSM_OUT_* flags)
    void initialize(int fftpoint, int batch, int in_format, int out_flags, int phase_mode = MSOSM_PHASE_FP64);
    void filter_block_int8(const int8_t *iq);
    void filter_block_c64(const Complex *x);
    void filter_block_device(const void *dev_input);
    void get_output(Complex *output);
    // fold every DM's power inside the dedispersion kernel (Fold_GPU built with num_series == numDMs)
    void attach_fold(Fold_GPU *fold);
    void *get_stream();
    // the stream get_output copies on (the reference's output_stream)
    void *get_output_stream();
    msosm_plan *plan() { return plan_; }
    int batch_size() const { return batch; }
    std::vector<Segment> build_segments(const int *delay_block, int fft_len);

protected:
    bool fold = false;
    int batch = 0;
    int fftpoint = 0;
    int in_format = MSOSM_IN_U16;
    int out_flags = MSOSM_OUT_U16;
    msosm_plan *plan_ = nullptr;
    void push(const void *input, int mode);
};
