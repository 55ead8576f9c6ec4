// Shared types of the reference-compatible C++ API (utils/include/globals.h:22-32).
#pragma once

#include <cmath>
#include <cstdint>
#include <iostream>
#include <string>

#include "msosm_amd.h"

using namespace std;

// offset-binary uint16 I/Q pair, code 0 reserved for 0 (globals.h:22-26)
struct __attribute__((packed)) uint16_pair {
    uint16_t first;
    uint16_t second;
};

// complex64, layout-compatible with cufftComplex / hipFloatComplex / float2 (globals.h:29)
struct Complex {
    float x;
    float y;
};

#ifndef FFTW3_H
typedef float fftwf_complex[2]; // same type FFTW declares, so both headers can be included together
#endif

extern bool verbose;
constexpr auto pi = M_PI;
void banner();
void line();

namespace msosm {
// Reference error convention: report and exit(1) (gpu/include/gpu_common.cuh:42-62). With
// set_throw_on_error(true) a std::runtime_error is thrown instead (tests, embedding).
void set_throw_on_error(bool on);
void check(int status, const char *what);
} // namespace msosm
