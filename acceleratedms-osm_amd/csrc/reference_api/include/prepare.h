// Prepare_OSM: host design of plain overlap-save (dedispersion/include/prepare.h:9-40,
// dedispersion/overlap_save/osm_prepare.cpp).
#pragma once

#include "globals.h"

class Prepare_OSM {
public:
    Prepare_OSM(float bw = 0, float dm = 0, float f0 = 0);
    virtual ~Prepare_OSM();
    void calculate_min_order();
    void override_order(unsigned long order);
    void generate_dedisp_params();

public:
    unsigned long M = 0;
    unsigned long M_min = 0;
    fftwf_complex *dedisp_params = nullptr;
    unsigned long Nd = 0;

protected:
    float bw;
    float fs;
    float kdm;
    float dm;
    float f0;
    float w0;
};
