// MSOSM_GPU_UNEVEN over the msosm_multi_* C ABI.
#include "msosm_gpu_uneven.h"

#include <vector>

MSOSM_GPU_UNEVEN::MSOSM_GPU_UNEVEN(float bw, float *dm, float f0, int n)
    : numDMs(n), dm_values(dm, dm + n), bw_(bw), f0_(f0)
{
    // per-DM design (Prepare_MSOSM's rules per DM, msosm_prepare.cpp:3-69)
    M_values.resize(n);
    delaycount_values.resize(n);
    for (int i = 0; i < n; ++i) {
        int nd = 0;
        msosm::check(msosm_design_nd(bw, f0, dm[i], &nd), "MSOSM_GPU_UNEVEN");
        msosm::check(msosm_design_uneven_order(nd, &M_values[i]), "MSOSM_GPU_UNEVEN");
        msosm::check(msosm_design_delaycount(nd, M_values[i], &delaycount_values[i]), "MSOSM_GPU_UNEVEN");
        if (delaycount_values[i] < 1) delaycount_values[i] = 1;
    }
}

MSOSM_GPU_UNEVEN::~MSOSM_GPU_UNEVEN()
{
    if (multi_) msosm_multi_destroy(multi_);
}

void MSOSM_GPU_UNEVEN::initialize(int push_blocks, int in_format, int out_flags, int phase_mode)
{
    if (multi_) msosm_multi_destroy(multi_);
    multi_ = nullptr;
    in_format_ = in_format;
    msosm_multi_desc d{};
    d.num_dms = numDMs;
    d.M = M_values.data();
    d.delaycount = delaycount_values.data();
    d.push_blocks = push_blocks;
    d.in_format = in_format;
    d.out_flags = out_flags;
    d.device = -1;
    d.stream = nullptr;
    msosm::check(msosm_multi_create(&d, &multi_), "initialize");
    msosm::check(msosm_multi_get_info(multi_, 0, nullptr, &classes_, &push_), "initialize");
    for (int i = 0; i < numDMs; ++i) {
        const int M = M_values[i];
        std::vector<int> delay(2 * (size_t)M);
        std::vector<float> H(4 * (size_t)M);
        int dc = 0;
        msosm::check(msosm_design_segmentation(bw_, f0_, dm_values[i], M, delay.data(), &dc), "initialize");
        msosm::check(msosm_design_transfer_function(bw_, f0_, dm_values[i], M, phase_mode, H.data()), "initialize");
        msosm::check(msosm_multi_set_dm(multi_, i, delay.data(), H.data()), "initialize");
    }
}

void MSOSM_GPU_UNEVEN::push(const void *input)
{
    if (!multi_) msosm::check(MSOSM_ERR_STATE, "filter_block: call initialize first");
    msosm::check(msosm_multi_push(multi_, input, MSOSM_INPUT_HOST_SYNC), "filter_block");
}

void MSOSM_GPU_UNEVEN::filter_block_uint16(uint16_pair *input)
{
    if (in_format_ != MSOSM_IN_U16) msosm::check(MSOSM_ERR_STATE, "filter_block_uint16: input format is not uint16");
    push(input);
}

void MSOSM_GPU_UNEVEN::filter_block_int8(const int8_t *iq)
{
    if (in_format_ != MSOSM_IN_I8) msosm::check(MSOSM_ERR_STATE, "filter_block_int8: input format is not int8");
    push(iq);
}

void MSOSM_GPU_UNEVEN::filter_block_c64(const Complex *x)
{
    if (in_format_ != MSOSM_IN_C64) msosm::check(MSOSM_ERR_STATE, "filter_block_c64: input format is not complex64");
    push(x);
}

void MSOSM_GPU_UNEVEN::filter_block_device(const void *dev_input)
{
    if (!multi_) msosm::check(MSOSM_ERR_STATE, "filter_block_device: call initialize first");
    msosm::check(msosm_multi_push(multi_, dev_input, MSOSM_INPUT_DEVICE), "filter_block_device");
}

void MSOSM_GPU_UNEVEN::get_output(int dm_index, Complex *output)
{
    msosm::check(msosm_multi_get_output(multi_, dm_index, output), "get_output");
}

Complex *MSOSM_GPU_UNEVEN::get_output_pointer(int dm_index)
{
    void *p = nullptr;
    msosm::check(msosm_multi_output_c64(multi_, dm_index, &p), "get_output_pointer");
    return static_cast<Complex *>(p);
}

void MSOSM_GPU_UNEVEN::fold(float period, unsigned long nbins)
{
    msosm::check(msosm_multi_fold(multi_, period, bw_, nbins), "fold");
    nbins_ = nbins;
}

std::vector<float> MSOSM_GPU_UNEVEN::get_folded_data()
{
    std::vector<float> prof((size_t)numDMs * nbins_);
    msosm::check(msosm_multi_fold_read(multi_, prof.data(), nullptr, nullptr), "get_folded_data");
    return prof;
}

void MSOSM_GPU_UNEVEN::set_launch_mode(int mode) { msosm::check(msosm_multi_set_launch_mode(multi_, mode), "set_launch_mode"); }

void MSOSM_GPU_UNEVEN::set_graph(bool on) { msosm::check(msosm_multi_set_graph(multi_, on ? 1 : 0), "set_graph"); }

void MSOSM_GPU_UNEVEN::synchronize() { msosm::check(msosm_multi_synchronize(multi_), "synchronize"); }
