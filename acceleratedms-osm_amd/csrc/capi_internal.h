// Internal declarations shared by the C-ABI implementation files.
#pragma once

#include <complex>
#include <string>

namespace msosm {

struct Fft64Handle;

int design_nd(float bw, float f0, float dm);
int design_min_order(int nd);
unsigned long design_osm_order(unsigned long nd);
int design_delaycount(int nd, int M);
int design_segmentation(float bw, float f0, float dm, int M, int *delay_points);
int design_build_segments(const int *delay, int fft_len, int max_len, int *seg, int cap);
int design_transfer_function(float bw, float f0, float dm, int M, int phase_mode, float *H, const Fft64Handle *fft,
                             int deriv = 0);
int design_transfer_functions(float bw, float f0, const float *dms, int num_dms, int M, int phase_mode, float *H,
                              int nthreads, int deriv = 0);
const Fft64Handle *design_fft_create(int M);
void design_fft_destroy(const Fft64Handle *h);
// unnormalised FP64 transform in place, sign -1: e^{-i}, +1: e^{+i}
void design_fft_run(const Fft64Handle *h, std::complex<double> *x, int sign);
// e^{+2 pi i k/n} in FP64 (octant-reduced)
void unit_root(long k, long n, double *re, double *im);

// thread-local last-error text; returns `code` for tail calls
int set_error(int code, const std::string &msg);

} // namespace msosm
