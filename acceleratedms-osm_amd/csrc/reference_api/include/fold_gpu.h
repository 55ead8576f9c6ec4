// Fold_GPU: phase folding of dedispersed power (folding/include/fold_gpu.h:10-43,
// folding/fold_gpu.cpp) on the HIP fold kernels.
#pragma once

#include <fstream>
#include <string>

#include "msosm_gpu_batch.h"
#include "osm_gpu_batch.h"

class Fold_GPU {
public:
    Fold_GPU(float period, float fs, unsigned long size, string outfileName = "fold.txt",
             unsigned long time_bin = (unsigned long)-1, int num_series = 1);
    // two polarisations: amplitude sqrt(|a|^2 + |b|^2); one: power |a|^2 (gpu_kernel.cu:430-444)
    template <typename T>
    void calculate_intensity(T *pol1, T *pol2);
    template <typename T>
    void calculate_intensity(T *pol1);
    void fold_data_phase();
    void get_folded_data();
    void write_to_file();
    ~Fold_GPU();

    msosm_fold *handle() { return fold_; }

public:
    float *folded_data;
    unsigned long time_bin;
    unsigned long period_samples;

private:
    unsigned long size;
    int num_series;
    long long series_stride = 0;
    const void *pol1_d = nullptr;
    const void *pol2_d = nullptr;
    void *stream = nullptr, *stream2 = nullptr;
    msosm_fold *fold_ = nullptr;
    ofstream outfile;
};
