/*
 * CPU oracle for the MS-OSM coherent-dedispersion path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference's algorithm, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product never links, loads or calls anything in this directory.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root). Pinning of each function (closed forms,
 * invariants, brute force, numpy) lives in tests/test_oracle_*.py.
 */
#ifndef MSOSM_ORACLE_H
#define MSOSM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (FP64, radix-2, power-of-two n). sign=-1: e^{-i} (forward), +1: e^{+i}. Unnormalised. */
int orc_fft(double *re_im, long n, int sign);

/* ---- design (reference-exact host arithmetic; compile with -ffp-contract=off) ---- */
/* Prepare_MSOSM ctor Nd: dedispersion/msosm/msosm_prepare.cpp:12,33 */
int orc_nd(float bw, float f0, float dm);
/* Prepare_MSOSM::calculate_min_order per-DM M_i: msosm_prepare.cpp:59-61 */
int orc_min_order(int nd);
/* Prepare_OSM::calculate_min_order: overlap_save/osm_prepare.cpp:17-30 (+ next_power_of_2 :127-140) */
long orc_osm_min_order(unsigned long nd);
/* delaycount for one DM: msosm_prepare.cpp:217 (Nd==0 is defined as 1: a documented fix of the
 * reference's uninitialised table for DM=0). */
int orc_delaycount(int nd, int M);
/* Prepare_MSOSM::segmentation(i): msosm_prepare.cpp:215-284. Writes delay_points[0..2M).
 * Uses the DM's own delaycount (identical to the reference for ascending DM lists).
 * Returns delaycount, or -1 on bad arguments. */
int orc_segmentation(float bw, float f0, float dm, int M, int *delay_points);
/* MSOSM_GPU_BATCH::build_segments: msosm_gpu_batch.cpp:262-295 (max_len = BLOCK_SIZE = 256).
 * Writes (start,len,group) triples into seg[3*cap]; returns the count or -1 if cap is too small. */
int orc_build_segments(const int *delay, int fft_len, int max_len, int *seg, int cap);
/* Prepare_MSOSM::generate_dedisp_params(i): msosm_prepare.cpp:107-164, restated in FP64
 * (FP64 phase, FP64 FFTs). H is 2M complex values, interleaved re/im doubles.
 * phase_fp32 != 0 selects the legacy FP32-phase form of msosm_prepare.cpp:166-213. */
int orc_transfer_function(float bw, float f0, float dm, int M, int phase_fp32, double *H);
/* dH/dDM of the same design (no reference counterpart): the design map is linear in the chirp, and
 * the phase is proportional to DM, so dc/dDM = -i (phase at DM 1) c. */
int orc_transfer_function_ddm(float bw, float f0, float dm, int M, int phase_fp32, double *dH);

/* ---- MS-OSM / OSM (FP64) ----
 * A6-A7 of the reference path: windows x[(g-1)M,(g+1)M) with x[<0]=0 (msosm_gpu_batch.cpp:47-52,
 * 153-176), X_g = DFT_2M, Y[k] = X_{g-delay(k)}[k] * H[k] with X_{<0}=0 (gpu_kernel.cu:139-189),
 * y = IDFT_2M(Y), out[gM+m] = y[M+m]/(2M) (gpu_kernel.cu:346-367).
 * x: n_samples complex (interleaved doubles), n_samples % M == 0.
 * delays: [num_dms][2M] ints, H: [num_dms][2M] complex doubles, out: [num_dms][n_samples] complex.
 * Everything before sample 0 (input history and older spectra) is zero. */
int orc_msosm(const double *x, long n_samples, int M, int num_dms, const int *delays, const double *H,
              double *out);

/* ---- fold (reference-exact bin arithmetic) ----
 * foldDataPhase_kernel: gpu/gpu_kernel.cu:462-496 and Fold_GPU::fold_data_phase folding/fold_gpu.cpp:48-54.
 * Accumulates intensity[i] (skipping exact zeros) into sums/counts; returns nothing.
 * bins_out (optional) receives the bin of every sample (or -1 for skipped zeros). */
void orc_fold(const double *intensity, long size, unsigned long nbins, double t_start, double dt,
              double period, double *sums, unsigned long long *counts, long *bins_out);
/* bin of sample i under the reference formula (no FMA contraction) */
long orc_fold_bin(long i, unsigned long nbins, double t_start, double dt, double period);

/* ---- codecs ---- */
/* uint16ToComplex_kernel: gpu_kernel.cu:16-30 */
void orc_unpack_u16(const uint32_t *packed, long n, double *out);
/* discardSamplesToUint16_kernel encoding: gpu_kernel.cu:407-412, with values clamped to 65535
 * (documented fix of the reference's overflow into the imaginary half). Input is the FP32 value. */
uint32_t orc_encode_u16_pair(float re, float im);

/* ---- seeded synthetic pulsar baseband (test input generator) ----
 * Gaussian pulse (FWHM = duty*P) of peak amplitude amp modulating unit-variance complex noise,
 * dispersed over the whole record with the direct chirp e^{+i phi} (the simulator's convention,
 * simulated_pulsar/simulated_complex_signal.cpp:171-231), plus complex AWGN of std sigma.
 * n must be a power of two. out: n complex (interleaved doubles). */
int orc_pulsar(long n, double fs, double f0, double dm, double period, double duty, double amp,
               double sigma, uint64_t seed, double *out);

#ifdef __cplusplus
}
#endif
#endif
