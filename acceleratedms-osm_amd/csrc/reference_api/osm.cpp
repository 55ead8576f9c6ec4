// Prepare_OSM and OSM_GPU_BATCH (dedispersion/overlap_save/osm_prepare.cpp, osm_gpu_batch.cpp).
#include <vector>

#include "osm_gpu_batch.h"

Prepare_OSM::Prepare_OSM(float bw, float dm, float f0) : bw(bw), dm(dm), f0(f0)
{
    // osm_prepare.cpp:3-15
    fs = bw;
    kdm = (float)(4.15 * 1e15);
    w0 = (float)(2 * pi * f0);
    int nd = 0;
    if (bw > 0 && f0 > 0) msosm::check(msosm_design_nd(bw, f0, dm, &nd), "Prepare_OSM");
    Nd = (unsigned long)nd;
    if (verbose) {
        cout << "Sampling frequency: " << fs << endl;
        cout << "Start frequency: " << f0 << endl;
        cout << "Dispersion smearing points: " << Nd << endl;
    }
}

Prepare_OSM::~Prepare_OSM() { delete[] dedisp_params; }

void Prepare_OSM::calculate_min_order()
{
    // osm_prepare.cpp:17-30
    msosm::check(msosm_design_osm_order(Nd, &M), "calculate_min_order");
    M_min = M;
    if (verbose) {
        cout << "Prepare parameters..." << endl;
        cout << "Minimum filter order: " << M_min << endl;
    }
}

void Prepare_OSM::override_order(unsigned long order) { M = order; }

void Prepare_OSM::generate_dedisp_params()
{
    // osm_prepare.cpp:68-125 (same transfer function as the MS-OSM design)
    delete[] dedisp_params;
    dedisp_params = new fftwf_complex[2 * M];
    msosm::check(msosm_design_transfer_function(bw, f0, dm, (int)M, MSOSM_PHASE_FP64,
                                                reinterpret_cast<float *>(dedisp_params)),
                 "generate_dedisp_params");
}

OSM_GPU_BATCH::OSM_GPU_BATCH(float bw, float dm, float f0) : Prepare_OSM(bw, dm, f0) {}

OSM_GPU_BATCH::~OSM_GPU_BATCH()
{
    if (plan_) msosm_plan_destroy(plan_);
}

void OSM_GPU_BATCH::get_device_info() { cout << "MS-OSM MI355X library version " << msosm_version() << endl; }

void OSM_GPU_BATCH::initialize_uint16(unsigned long fftpoint, int batch, bool streaming)
{
    // osm_gpu_batch.cpp:19-92
    this->streaming = streaming;
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
    generate_dedisp_params();
    if (plan_) msosm_plan_destroy(plan_);
    msosm_plan_desc d{};
    d.M = (int)M;
    d.num_dms = 1;
    d.batch = batch;
    d.delaycount = 1;
    d.in_format = MSOSM_IN_U16;
    d.out_flags = MSOSM_OUT_C64;
    d.device = -1;
    msosm::check(msosm_plan_create(&d, &plan_), "initialize_uint16");
    std::vector<int> no_delay(2 * M, 0);
    msosm::check(msosm_plan_set_dm(plan_, 0, no_delay.data(), reinterpret_cast<const float *>(dedisp_params)),
                 "initialize_uint16");
    msosm_plan_info info{};
    msosm::check(msosm_plan_get_info(plan_, &info), "initialize_uint16");
    total_memory = info.device_bytes;
}

void OSM_GPU_BATCH::filter_block_uint16(uint16_pair *input)
{
    // osm_gpu_batch.cpp:94-123: synchronous copy unless streaming
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "filter_block_uint16: call initialize_uint16 first");
    msosm::check(msosm_plan_push(plan_, input, streaming ? MSOSM_INPUT_HOST_ASYNC : MSOSM_INPUT_HOST_SYNC),
                 "filter_block_uint16");
}

void OSM_GPU_BATCH::get_output(Complex *output)
{
    // osm_gpu_batch.cpp:125-133
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "get_output: not initialised");
    msosm::check(msosm_plan_get_output(plan_, MSOSM_OUT_C64, output), "get_output");
    if (!streaming) synchronize();
}

Complex *OSM_GPU_BATCH::get_output_pointer()
{
    void *p = nullptr;
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "get_output_pointer: not initialised");
    msosm::check(msosm_plan_output_c64(plan_, &p), "get_output_pointer");
    return static_cast<Complex *>(p);
}

void *OSM_GPU_BATCH::get_stream()
{
    void *s = nullptr;
    if (plan_) msosm::check(msosm_plan_get_stream(plan_, &s), "get_stream");
    return s;
}

void OSM_GPU_BATCH::synchronize()
{
    if (plan_) msosm::check(msosm_plan_synchronize(plan_), "synchronize");
}
