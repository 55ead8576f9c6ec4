// MSOSM_GPU_BATCH (dedispersion/msosm/msosm_gpu_batch.cpp:3-295) over the device plan of the C ABI.
#include "msosm_gpu_batch.h"

#include <cstdlib>
#include <thread>

#include "fold_gpu.h"

MSOSM_GPU_BATCH::MSOSM_GPU_BATCH(float bw, float dm, float f0) : Prepare_MSOSM(bw, dm, f0) {}

MSOSM_GPU_BATCH::MSOSM_GPU_BATCH(float bw, float *dm, float f0, int numDMs) : Prepare_MSOSM(bw, dm, f0, numDMs) {}

MSOSM_GPU_BATCH::~MSOSM_GPU_BATCH() { reset_device(); }

void MSOSM_GPU_BATCH::get_device_info()
{
    cout << "MS-OSM MI355X library version " << msosm_version() << endl;
}

void MSOSM_GPU_BATCH::initialize(int fftpoint, int batch, int in_format, int out_flags, int phase_mode)
{
    // msosm_gpu_batch.cpp:12-151
    if (plan_) reset_device();
    this->phase_mode = phase_mode;
    this->in_format = in_format;
    this->out_flags = out_flags;
    if (verbose) get_device_info();
    calculate_min_order();
    if (fftpoint != 0) {
        M = fftpoint / 2;
        if (M < M_min) {
            cout << "Warning: The filter order is too small for the given FFT point. The minimum FFT point "
                 << 2 * M_min << " is used." << endl;
            M = M_min;
        }
    }
    this->fftpoint = 2 * M;
    this->batch = batch;
    delaycount = 0;
    for (int i = 0; i < numDMs; i++) segmentation(i);
    // transfer functions for all DMs on all host cores (generate_dedisp_params(i), batched)
    float *hblock = new float[4 * (size_t)M * numDMs];
    msosm::check(msosm_design_transfer_functions(bw, f0, dm_values, numDMs, M, phase_mode, hblock, 0),
                 "generate_dedisp_params");
    for (int i = 0; i < numDMs; i++) {
        delete[] dedisp_params[i];
        dedisp_params[i] = new fftwf_complex[2 * (size_t)M];
        std::copy(hblock + 4 * (size_t)M * i, hblock + 4 * (size_t)M * (i + 1), &dedisp_params[i][0][0]);
    }
    delete[] hblock;

    msosm_plan_desc d{};
    d.M = M;
    d.num_dms = numDMs;
    d.batch = batch;
    d.delaycount = delaycount;
    d.in_format = in_format;
    d.out_flags = out_flags;
    d.device = -1;
    d.stream = nullptr;
    msosm::check(msosm_plan_create(&d, &plan_), "initialize");
    for (int i = 0; i < numDMs; i++)
        msosm::check(msosm_plan_set_dm(plan_, i, delay_points[i], reinterpret_cast<const float *>(dedisp_params[i])),
                     "initialize");
    cout << "FFT point: " << this->fftpoint << endl;
}

void MSOSM_GPU_BATCH::initialize_uint16(int fftpoint, int batch, bool fold)
{
    if (batch % ITEMS_PER_THREAD != 0) {
        // msosm_gpu_batch.cpp:130-134
        cout << "Batch size must be a multiple of " << ITEMS_PER_THREAD << " for the current implementation." << endl;
        exit(1);
    }
    this->fold = fold;
    initialize(fftpoint, batch, MSOSM_IN_U16, fold ? MSOSM_OUT_C64 : MSOSM_OUT_U16);
}

void MSOSM_GPU_BATCH::push(const void *input, int mode)
{
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "filter_block: call initialize_uint16 first");
    msosm::check(msosm_plan_push(plan_, input, mode), "filter_block");
}

void MSOSM_GPU_BATCH::filter_block_uint16(uint16_pair *input)
{
    // fold mode copies synchronously (msosm_gpu_batch.cpp:183); save mode is the async stream
    // pipeline (:207) and expects pinned input
    if (in_format != MSOSM_IN_U16) msosm::check(MSOSM_ERR_STATE, "filter_block_uint16: plan expects another format");
    push(input, fold ? MSOSM_INPUT_HOST_SYNC : MSOSM_INPUT_HOST_ASYNC);
}

void MSOSM_GPU_BATCH::filter_block_int8(const int8_t *iq)
{
    if (in_format != MSOSM_IN_I8) msosm::check(MSOSM_ERR_STATE, "filter_block_int8: plan expects another format");
    push(iq, MSOSM_INPUT_HOST_SYNC);
}

void MSOSM_GPU_BATCH::filter_block_c64(const Complex *x)
{
    if (in_format != MSOSM_IN_C64) msosm::check(MSOSM_ERR_STATE, "filter_block_c64: plan expects another format");
    push(x, MSOSM_INPUT_HOST_SYNC);
}

void MSOSM_GPU_BATCH::filter_block_device(const void *dev_input) { push(dev_input, MSOSM_INPUT_DEVICE); }

void MSOSM_GPU_BATCH::get_output(uint16_pair *output)
{
    // msosm_gpu_batch.cpp:238-245 (asynchronous; synchronize() before reading)
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "get_output: not initialised");
    msosm::check(msosm_plan_get_output(plan_, MSOSM_OUT_U16, output), "get_output");
}

void MSOSM_GPU_BATCH::get_output(Complex *output)
{
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "get_output: not initialised");
    msosm::check(msosm_plan_get_output(plan_, MSOSM_OUT_C64, output), "get_output");
}

Complex *MSOSM_GPU_BATCH::get_output_pointer()
{
    void *p = nullptr;
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "get_output_pointer: not initialised");
    msosm::check(msosm_plan_output_c64(plan_, &p), "get_output_pointer");
    return static_cast<Complex *>(p);
}

void *MSOSM_GPU_BATCH::get_stream()
{
    void *s = nullptr;
    if (plan_) msosm::check(msosm_plan_get_stream(plan_, &s), "get_stream");
    return s;
}

void *MSOSM_GPU_BATCH::get_output_stream()
{
    void *s = nullptr;
    if (plan_) msosm::check(msosm_plan_get_output_stream(plan_, &s), "get_output_stream");
    return s;
}

void MSOSM_GPU_BATCH::attach_fold(Fold_GPU *f)
{
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "attach_fold: not initialised");
    msosm::check(msosm_plan_attach_fold(plan_, f ? f->handle() : nullptr), "attach_fold");
}

void MSOSM_GPU_BATCH::synchronize()
{
    if (plan_) msosm::check(msosm_plan_synchronize(plan_), "synchronize");
}

void MSOSM_GPU_BATCH::reset_device()
{
    if (plan_) {
        msosm_plan_destroy(plan_);
        plan_ = nullptr;
    }
}

std::vector<Segment> MSOSM_GPU_BATCH::build_segments(const int *delay_block, int fft_len)
{
    // msosm_gpu_batch.cpp:262-295
    std::vector<int> seg(3 * (size_t)fft_len + 3);
    int count = 0;
    msosm::check(msosm_design_build_segments(delay_block, fft_len, BLOCK_SIZE, seg.data(), fft_len + 1, &count),
                 "build_segments");
    std::vector<Segment> out(count);
    for (int i = 0; i < count; ++i) out[i] = Segment{seg[3 * i], seg[3 * i + 1], seg[3 * i + 2]};
    return out;
}
