// Fold_GPU (folding/fold_gpu.cpp:4-77) over the fold object of the C ABI.
#include "fold_gpu.h"

Fold_GPU::Fold_GPU(float period, float fs, unsigned long size, string outfileName, unsigned long time_bin,
                   int num_series)
    : size(size), num_series(num_series)
{
    // fold_gpu.cpp:4-25; the bin arithmetic runs in double on the device (t_start, dt = 1/fs, period)
    const float period_samples_float = period * fs;
    period_samples = static_cast<unsigned long>(period_samples_float);
    if (time_bin == (unsigned long)-1) time_bin = period_samples;
    this->time_bin = time_bin;
    folded_data = new float[time_bin * (size_t)num_series]();
    outfile.open(outfileName);
    msosm::check(msosm_fold_create((double)period, (double)fs, size, time_bin, num_series, -1, &fold_), "Fold_GPU");
    series_stride = (long long)size;
}

template <typename T>
void Fold_GPU::calculate_intensity(T *pol1, T *pol2)
{
    // fold_gpu.cpp:27-34; the power is computed inside the fold kernel
    pol1_d = pol1->get_output_pointer();
    pol2_d = pol2->get_output_pointer();
    stream = pol1->get_stream();
    stream2 = pol2->get_stream();
}

template <typename T>
void Fold_GPU::calculate_intensity(T *pol1)
{
    // fold_gpu.cpp:36-41
    pol1_d = pol1->get_output_pointer();
    pol2_d = nullptr;
    stream = pol1->get_stream();
    stream2 = stream;
}

template void Fold_GPU::calculate_intensity<MSOSM_GPU_BATCH>(MSOSM_GPU_BATCH *, MSOSM_GPU_BATCH *);
template void Fold_GPU::calculate_intensity<OSM_GPU_BATCH>(OSM_GPU_BATCH *, OSM_GPU_BATCH *);
template void Fold_GPU::calculate_intensity<MSOSM_GPU_BATCH>(MSOSM_GPU_BATCH *);
template void Fold_GPU::calculate_intensity<OSM_GPU_BATCH>(OSM_GPU_BATCH *);

void Fold_GPU::fold_data_phase()
{
    // fold_gpu.cpp:48-54 (t_start advances by size * dt inside the fold object)
    if (!pol1_d) msosm::check(MSOSM_ERR_STATE, "fold_data_phase: call calculate_intensity first");
    // each object dedisperses on its own stream: the fold (on pol1's) waits for pol2's output too
    if (pol2_d) msosm::check(msosm_stream_wait(stream, stream2), "fold_data_phase");
    msosm::check(msosm_fold_accumulate(fold_, pol1_d, pol2_d, series_stride, stream), "fold_data_phase");
}

void Fold_GPU::get_folded_data()
{
    // fold_gpu.cpp:56-60
    msosm::check(msosm_fold_read(fold_, folded_data, nullptr, nullptr), "get_folded_data");
}

void Fold_GPU::write_to_file()
{
    // fold_gpu.cpp:62-68
    for (unsigned long i = 0; i < time_bin * (unsigned long)num_series; i++) outfile << folded_data[i] << endl;
}

Fold_GPU::~Fold_GPU()
{
    outfile.close();
    if (fold_) msosm_fold_destroy(fold_);
    delete[] folded_data;
}
