// Prepare_MSOSM: host design of the multi-segment overlap-save method
// (dedispersion/include/msosm_prepare.h:9-60, dedispersion/msosm/msosm_prepare.cpp).
#pragma once

#include "globals.h"

constexpr int MIN_SEGMENT_POINTS = 1024;

class Prepare_MSOSM {
public:
    Prepare_MSOSM(float bw = 0, float dm = 0, float f0 = 0);
    Prepare_MSOSM(float bw, float *dm, float f0, int numDMs = 1);
    virtual ~Prepare_MSOSM();
    void calculate_min_order();
    void override_order(int order);
    // H of DM i into dedisp_params[i] (2M complex); FP64 phase and FP64 transforms, rounded once
    void generate_dedisp_params(int i = 0);
    // per-bin delays of DM i into delay_points[i] (2M ints)
    void segmentation(int i = 0);

public:
    int numDMs;
    int M = 0;
    int M_min = 0;
    int *M_values = nullptr;
    int delaycount = 0;
    int *delaycount_values = nullptr;
    fftwf_complex **dedisp_params = nullptr;
    int **delay_points = nullptr;
    int *Nd_values = nullptr;
    int start_Nd = 0;
    int end_Nd = 0;

protected:
    float bw;
    float fs;
    float kdm;
    float *dm_values;
    float f0;
    float w0;
    int phase_mode = MSOSM_PHASE_FP64;

private:
    bool own_dm_values = false;
    int designed_M = 0; // M the per-DM tables were built for
    void release_tables();
};
