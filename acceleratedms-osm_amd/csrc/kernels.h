// Kernel argument blocks and launchers of the MS-OSM HIP path (device code in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "msosm_amd.h"

namespace msosm {

struct FwdArgs {
    const void *in;        // this push's B*M input samples (device)
    int in_fmt;            // MSOSM_IN_*
    const float2 *hist;    // the M samples preceding `in` as complex64 (zeros before the stream start)
    float2 *ring;          // [FBS][L] spectra
    int FBS;
    long long g0;          // stream block index of this push's first block
    int B;                 // blocks per push
    float2 *scratch;       // [grid][L]: a slot per workgroup (fwd_kernel) or per block (fwd_split_kernel)
    const float2 *tw_n1, *tw_n2, *tw_lo, *tw_hi;
    int tw_s;
    float2 *ring_p;        // [FBS][L] the same spectra in the accumulate layout (perm_pos), or null
    int perm_rl, perm_sl;  //   its bin k at ((k mod 2^rl) << sl) | (k >> rl)
};

// Position of spectrum bin k in the accumulate layout of a plan's ring and chirp tables:
// P(k) = (k mod R) S + k div R with R = 2^rl, S = 2^sl (the accumulate launch's decimation in time;
// rl = 0 is the natural order).
__host__ __device__ __forceinline__ uint32_t perm_pos(uint32_t k, int rl, int sl)
{
    return ((k & ((1u << rl) - 1u)) << sl) | (k >> rl);
}


// Index of accumulate-layout position P = r S + p + 1024 t (S = 2^14; p in [0, 1024) the column index
// of the transform's thread tid = 256 (p mod 128 div 32) + p mod 32 + 32 (p div 128)) in the accumulate
// launch's chirp table: values t = 2i, 2i + 1 of thread tid form the 16-byte word of lane tid mod 64
// in row i of its wave's block, i.e. index r S + 1024 (tid div 64) + 128 i + 2 (tid mod 64) + (t mod 2),
// so that each of a wave's eight chirp loads reads 1 KiB in one run.
__host__ __device__ __forceinline__ uint32_t acc_chirp_pos(uint32_t P)
{
    const uint32_t r = P >> 14, q = P & 16383u, p = q & 1023u, t = q >> 10;
    const uint32_t tid = ((p & 127u) >> 5) * 256u + (p & 31u) + 32u * (p >> 7);
    return (r << 14) + (tid >> 6) * 1024u + (t >> 1) * 128u + (tid & 63u) * 2u + (t & 1u);
}

struct InvArgs {
    const float2 *ring;
    int FBS;
    long long g0;
    int B, D;
    const float2 *H;          // [D][L] transfer functions
    const uint16_t *delay;    // [D][L] per-bin delay in blocks
    float2 *scratch;          // [grid][L]
    float2 *out_c64;          // [D][B][M] or null
    uint32_t *out_u16;        // [D][B][M] or null
    uint32_t *out_f16;        // [D][B][M] FP16 pairs or null
    float *out_scale;         // [D][B] scales of out_f16
    const float2 *tangent_ref; // [D][B][M] primal series: fold 2 Re(conj(ref) y) instead of |y|^2, or null
    const uint16_t *bins;     // [B*M] fold bin of each call-relative sample, or null
    double *fold_sum;         // [D][nbins] or null (no fold)
    unsigned long long *fold_cnt;
    int nbins;
    const float2 *tw_n1, *tw_n2, *tw_lo, *tw_hi;
    int tw_s;
    int order;                // unit order (MSOSM_ORDER_*): which units run together
    // accumulate launch (lengths 2^14..2^16): the ring in the layout perm_pos with R = L / 2^14, H with
    // bin k at acc_chirp_pos(perm_pos(k)), and the delays as bytes in thread order: the 16 delays a
    // thread gathers for transform r are the 16 bytes at r 2^14 + 16 p (p = its column index in
    // [0, 1024)), the t-th one being bin position r 2^14 + p + 1024 t
    const float2 *ring_p, *H_p;
    const uint8_t *delay_acc;
    // sub-unit accumulate launch (lengths 2^17, 2^18: Q = L / 2^16 sub-units of 2^16 points per unit):
    // the twiddle tables of the sub-unit length (lo, hi split at sub_s; the 128-point roots), and the
    // sub-unit spectra [D][B][Q][2^16] that the gathering pass writes (sub-unit c of unit (d, b) at
    // ((d B + b) Q + c) 2^16, bin k' at acc_chirp_pos(perm_pos(k', 2, 14)): the chirp table's order)
    const float2 *sub_lo, *sub_hi, *sub_128;
    int sub_s;
    float2 *xpre;
    // per unit (d B + b): arrival counters of its sub-units, zeroed by the gathering pass (or null). The Q
    // sub-units of a unit wait for each other, up to a bounded number of polls, before writing their
    // samples, so that every cache line of the unit's series is filled while it sits in one L2.
    unsigned int *sub_sync;
};

// One persistent launch over several FFT lengths ("length classes"): class c owns the units
// [ustart[c], ustart[c+1]) (unit = (DM, block) of that class, block fastest); workgroups take units
// in increasing order from a device counter, so the long transforms (class 0 = largest) start first.
constexpr int MAX_CLASSES = 8;
struct MultiInvArgs {
    InvArgs cls[MAX_CLASSES];  // per class, as for one-length launches (rings below 4 GiB)
    int log2L[MAX_CLASSES];
    long long ustart[MAX_CLASSES + 1];
    int ncls;
    int nbins;                 // common fold bins of every class (0 = no fold)
    float2 *scratch;           // [grid][scr_stride] workgroup scratch slots
    long long scr_stride;      // >= the largest L
    unsigned int *counter;     // unit counter, zero at launch
};
// Batched transforms of the GPU pulsar simulator (kernels_sim.hip): transform t reads element e at
// src[t * in_t + e * in_e]; its output o goes to index di = t * out_t + o * out_e through the stage
// `mode`.
enum { SIM_PLAIN = 0, SIM_TWIDDLE = 1, SIM_CHIRP = 2, SIM_EXTRACT = 3 };
struct SimFftArgs {
    const float2 *src;
    float2 *dst;
    long long in_t, in_e, out_t, out_e;
    long long ntrans;
    int mode;
    long long big_n;        // SIM_TWIDDLE: dst[di] = y * e^{DIR 2 pi i (o t mod big_n) / big_n}
    const float2 *chirp;    // SIM_CHIRP: dst[di] = y * chirp[di]
    long long ex_start;     // SIM_EXTRACT: for ex_start <= di < ex_start + ex_count, v = y * ex_scale
    long long ex_count;     //   goes to ex_c64[di - ex_start] and, as the reference's uint16 pair,
    float ex_scale;         //   to ex_u16[di - ex_start] (either may be null)
    float2 *ex_c64;
    uint32_t *ex_u16;
    float2 *scratch;        // [grid][L]
    const float2 *tw_n1, *tw_n2, *tw_lo, *tw_hi;
    int tw_s;
};
bool sim_length_supported(int log2L);
hipError_t launch_sim_fft(int log2L, int dir, const SimFftArgs &a, int grid, hipStream_t st);

// A kernel launch as a graph node: function, launch shape and dynamic LDS (the arguments are the
// launcher's, in order).
struct NodeDesc {
    const void *func = nullptr;
    dim3 grid, block;
    unsigned int lds = 0;
};
NodeDesc fwd_node(int log2L, int grid);        // fwd_kernel(FwdArgs)
NodeDesc hist_node(int M);                     // hist_kernel(hist, in, fmt, first, M)
NodeDesc bins_node(long long n);               // bins_kernel(bins, n, nbins, t_start, dt, period)
NodeDesc fill_node(long long nbytes);          // fill_kernel(words, nwords, value)
NodeDesc inv_multi_node(const MultiInvArgs &a, int grid); // inv_multi_kernel(MultiInvArgs)

bool multi_length_supported(int log2L);
int multi_tiles_lds_bytes();
hipError_t configure_multi_lds(int bytes);
hipError_t launch_inv_multi(const MultiInvArgs &a, int grid, hipStream_t st);

void split_length(int log2L, int *n1, int *n2);
bool length_supported(int log2L);
bool split_supported(int log2L); // a role-split dedispersion kernel exists for this length
// accumulate form (kernels_accum.hip): lengths 2^14..2^16 (spectrum layout R = L / 2^14), and 2^17, 2^18
// as Q = accum_subunits sub-units of 2^16 points each (a gathering pass, then the sub-unit transforms;
// the plain and uint16/complex64/fold epilogues)
bool accum_supported(int log2L);
int accum_subunits(int log2L);
int accum_lds_bytes(int log2L);
hipError_t configure_accum_lds(int log2L, int bytes);
hipError_t launch_accum(int log2L, const InvArgs &a, int grid, hipStream_t st);
int tiles_lds_bytes(int log2L); // tiles and LDS twiddle tables of the forward and dedispersion kernels
// LDS the inverse kernel needs beyond its tiles and fold histogram (FP16 scale reduction)
constexpr int INV_EXTRA_LDS = 16;
int workgroup_threads();
hipError_t configure_lds(int log2L, int fwd_bytes, int inv_bytes);
// grid: fwd_kernel's workgroups (one scratch slot each); cus: workgroups the GPU holds at once, the
// bound of the split form's B P workgroups (used when grid = B)
hipError_t launch_fwd(int log2L, const FwdArgs &a, int grid, int cus, hipStream_t st);
hipError_t launch_inv(int log2L, const InvArgs &a, int grid, hipStream_t st);
hipError_t launch_hist(float2 *hist, const void *in, int fmt, long long first, int M, hipStream_t st);
hipError_t launch_bins(uint16_t *bins, long long n, unsigned long nbins, double t_start, double dt, double period,
                       hipStream_t st);
// the bins of a sub-unit accumulate launch: sample m = Q j + c of block b at (b Q + c) (M / Q) + j
hipError_t launch_bins_sub(uint16_t *bins, long long n, unsigned long nbins, double t_start, double dt, double period,
                           int q, long long m, hipStream_t st);

hipError_t launch_fold_series(const float2 *a, const float2 *b, long long size, long long stride, int nseries,
                              unsigned long nbins, double t_start, double dt, double period, double *sums,
                              unsigned long long *cnts, hipStream_t st);
// profile = sum / count (0 where empty); with cnt_out, the counts are copied there too
hipError_t launch_profile(const double *sums, const unsigned long long *cnts, float *prof, long long n,
                          hipStream_t st, unsigned long long *cnt_out = nullptr);
hipError_t launch_fill(void *p, long long nbytes, uint32_t value, hipStream_t st);

} // namespace msosm
