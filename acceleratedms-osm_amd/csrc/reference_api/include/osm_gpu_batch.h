// OSM_GPU_BATCH: plain overlap-save on the GPU (dedispersion/include/osm_gpu_batch.h:7-38,
// dedispersion/overlap_save/osm_gpu_batch.cpp). It runs on the MS-OSM kernels with one DM and no
// block delays, which is exactly overlap-save with an M-tap filter.
#pragma once

#include "prepare.h"

class OSM_GPU_BATCH : public Prepare_OSM {
public:
    OSM_GPU_BATCH(float bw = 0, float dm = 0, float f0 = 0);
    ~OSM_GPU_BATCH() override;
    void get_device_info();
    void initialize_uint16(unsigned long fftpoint = 0, int batch = 1, bool streaming = false);
    void filter_block_uint16(uint16_pair *input);
    void get_output(Complex *output);
    Complex *get_output_pointer();
    void *get_stream();
    void synchronize();

public:
    int batch = 0;
    unsigned long total_memory = 0;

private:
    bool streaming = false;
    unsigned long fftpoint = 0;
    msosm_plan *plan_ = nullptr;
};
