// HIP kernels of the MS-OSM coherent-dedispersion path for gfx950: the forward transform into the
// spectrum ring, the persistent dedispersion kernel of one FFT length, the small helper kernels and
// their launchers. The dedispersion kernel is persistent: one workgroup per CU walks (DM, block)
// units, with the unit order chosen so that the workgroups sharing an XCD (and its L2) work on the
// same DM's chirp at the same time.
#include "msosm_device.h"

namespace msosm {

// ---------------------------------------------------------------------------------------------
// Forward: unpack + window + DFT_L into ring slots. One unit = one block of M new samples.
// ---------------------------------------------------------------------------------------------
// Pass A of block b's transform (tiles [t_begin, t_end)): the block's L input samples (the M before it
// from the history at the stream start) into the block's scratch slot.
template <int N1, int N2>
__device__ __forceinline__ void fwd_block_a(const FwdArgs &a, int b, float2 *scr, float2 *lds_all, const Tw &tw,
                                            int t_begin, int t_end)
{
    constexpr int M = Shape<N1, N2>::L / 2;
    const long long base = (long long)(b - 1) * M;
    auto load = [&](int k) -> float2 {
        const long long s = base + k;
        if (s < 0) return a.hist[s + M];
        return decode_in(a.in, a.in_fmt, s);
    };
    pass_a<N1, N2, -1>(load, scr, lds_all, tw, t_begin, t_end);
}

// Pass B of block b's transform (tiles [t_begin, t_end)): from the scratch slot into the block's ring
// slot (and the permuted ring, when the plan keeps one).
template <int N1, int N2>
__device__ __forceinline__ void fwd_block_b(const FwdArgs &a, int b, const float2 *scr, float2 *lds_all,
                                            const Tw &tw, int t_begin, int t_end)
{
    using S = Shape<N1, N2>;
    constexpr int L = S::L;
    const long long slot = (a.g0 + b) % a.FBS;
    float2 *dst = a.ring + (size_t)slot * L;
    float2 *dstp = a.ring_p ? a.ring_p + (size_t)slot * L : nullptr;
    auto prep = [](int, bool) { return 0; };
    auto emit = [&](const float2 *lds, int tile, bool active, int) {
        if (!active) return;
        constexpr int NC = S::NCB;
        const int lt = threadIdx.x & 255;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int f = lt + 256 * i;
            const int n1 = f / NC, row = f % NC;
            const uint32_t k = (uint32_t)(tile * NC + row + N2 * n1);
            const float2 y = lds[lds_pad(lt) + 272 * i]; // pad(lt + 256 i)
            st32(dst, k, y);
            if (dstp) st32(dstp, perm_pos(k, a.perm_rl, a.perm_sl), y);
        }
    };
    pass_b<N1, N2, -1>(scr, prep, emit, lds_all, tw, 0, t_begin, t_end);
}

// One workgroup per block (blocks b, b + grid, ...), both passes through the workgroup's scratch slot.
template <int N1, int N2>
__global__ void __launch_bounds__(WG) fwd_kernel(FwdArgs a)
{
    using S = Shape<N1, N2>;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N1, N2>(lds_all, tw);
    float2 *scr = a.scratch + (size_t)blockIdx.x * S::L;
#pragma unroll 1
    for (int b = blockIdx.x; b < a.B; b += gridDim.x) {
        fwd_block_a<N1, N2>(a, b, scr, lds_all, tw, 0, S::NTA);
        __syncthreads();
        fwd_block_b<N1, N2>(a, b, scr, lds_all, tw, 0, S::NTB);
        __syncthreads();
    }
}

// The same transform with each block's passes split over P workgroups (a push of fewer blocks than
// the GPU has CUs): workgroup (b, p) = (blockIdx.x / P, blockIdx.x mod P) runs its P-th share of the
// tiles of pass PASS through block b's own scratch slot (the plan has one per block). Two launches,
// so that every pass-A tile of a block is in its slot before any pass-B tile reads it; the tiles and
// their arithmetic are fwd_kernel's, so the spectra are the same bits.
template <int N1, int N2, int PASS>
__global__ void __launch_bounds__(WG) fwd_split_kernel(FwdArgs a, int P)
{
    using S = Shape<N1, N2>;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N1, N2>(lds_all, tw);
    const int b = (int)blockIdx.x / P, p = (int)blockIdx.x % P;
    float2 *scr = a.scratch + (size_t)b * S::L;
    constexpr int NT = PASS == 0 ? S::NTA : S::NTB;
    const int share = NT / P;
    if constexpr (PASS == 0)
        fwd_block_a<N1, N2>(a, b, scr, lds_all, tw, p * share, (p + 1) * share);
    else
        fwd_block_b<N1, N2>(a, b, scr, lds_all, tw, p * share, (p + 1) * share);
}

// ---------------------------------------------------------------------------------------------
// Dedispersion of one FFT length
// ---------------------------------------------------------------------------------------------
template <int N1, int N2, bool RING32, int VAR>
__global__ void __launch_bounds__(WG) inv_kernel(InvArgs a)
{
    using S = Shape<N1, N2>;
    // scratch prefetch where its registers fit without spills (plain epilogue, L <= 2^16); delay
    // prefetch for every plain length
    constexpr bool PREFETCH = VAR == VAR_PLAIN && S::L <= (1 << 16);
    constexpr bool PREFETCH_DELAYS = VAR == VAR_PLAIN;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N1, N2>(lds_all, tw);
    float2 *scr = a.scratch + (size_t)blockIdx.x * S::L;
    double *hsum = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_all) + S::LDS_BYTES);
    unsigned int *hcnt = reinterpret_cast<unsigned int *>(hsum + a.nbins);
    float *pw_lds = reinterpret_cast<float *>(hcnt + a.nbins + (a.nbins & 1));
    fold_lds_init(hsum, hcnt, pw_lds, a.fold_sum ? a.nbins : 0);
    const UnitOrder o(a, gridDim.x, blockIdx.x);
#pragma unroll 1
    for (int s = o.first; s < o.end; s += o.step) {
        int d, b;
        if (o.at(s, a, d, b)) // uniform over the workgroup
            inv_unit<N1, N2, RING32, VAR, PREFETCH, SUBT, PREFETCH_DELAYS>(a, d, b, lds_all, tw, scr, hsum, hcnt,
                                                                           pw_lds);
    }
}

// Role-split form of the plain dedispersion kernel for square lengths (see inv_rs).
template <int N, bool RING32>
__global__ void __launch_bounds__(WG) inv_rs_kernel(InvArgs a)
{
    using S = Shape<N, N>;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N, N>(lds_all, tw);
    float2 *scr = a.scratch + (size_t)blockIdx.x * S::L;
    double *hsum = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_all) + S::LDS_BYTES);
    unsigned int *hcnt = reinterpret_cast<unsigned int *>(hsum + a.nbins);
    float *pw_lds = reinterpret_cast<float *>(hcnt + a.nbins + (a.nbins & 1));
    fold_lds_init(hsum, hcnt, pw_lds, a.fold_sum ? a.nbins : 0);
    inv_rs<N, RING32>(a, lds_all, tw, scr, hsum, hcnt);
}

// lengths with a role-split kernel: square, at least two tile pairs per pass, and rows of at most
// 256 points (at 512 x 512 the split is correct but slower than the sequential form: 8.1-8.4
// against 7.7-7.9 ms per launch)
template <int N1, int N2>
constexpr bool has_rs()
{
    return N1 == N2 && (N1 == 128 || N1 == 256);
}

// ---------------------------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------------------------
__global__ void hist_kernel(float2 *hist, const void *in, int fmt, long long first, int M)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) hist[i] = decode_in(in, fmt, first + i);
}

// Fold bin of call-relative sample i, exactly as the reference evaluates it (no contraction):
// t = t_start + (double)i * dt; phase = fmod(t / P, 1); bin = min(floor(phase * nbins), nbins-1)
// (gpu/gpu_kernel.cu:477-481).
__device__ __forceinline__ unsigned long fold_bin(long long i, unsigned long nbins, double t_start, double dt,
                                                  double period)
{
#pragma clang fp contract(off)
    const double t = t_start + (double)i * dt;
    const double x = t / period;
    const double phase = fmod(x, 1.0);
    unsigned long bin = (unsigned long)floor(phase * (double)nbins);
    if (bin >= nbins) bin = nbins - 1;
    return bin;
}

__global__ void bins_kernel(uint16_t *bins, long long n, unsigned long nbins, double t_start, double dt,
                            double period)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bins[i] = (uint16_t)fold_bin(i, nbins, t_start, dt, period);
}

// The same bins in the order of a sub-unit accumulate launch: sample m = Q j + c of block b (M samples
// per block) at (b Q + c) (M / Q) + j, so that each sub-unit c reads its samples' bins contiguously
__global__ void bins_sub_kernel(uint16_t *bins, long long n, unsigned long nbins, double t_start, double dt,
                                double period, int q, long long m)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long b = i / m, mm = i - b * m, j = mm / q, c = mm - j * q;
    bins[b * m + c * (m / q) + j] = (uint16_t)fold_bin(i, nbins, t_start, dt, period);
}

// Unfused fold of series (Fold_GPU::calculate_intensity + fold_data_phase, fold_gpu.cpp:27-54):
// intensity |a|^2 (one polarisation) or sqrt(|a|^2+|b|^2) (two), FP64 sums and u64 counts.
// nseries series of `size` samples each, series s at a + s*stride.
__global__ void fold_series_kernel(const float2 *a, const float2 *b, long long size, long long stride, int nseries,
                                   unsigned long nbins, double t_start, double dt, double period, double *sums,
                                   unsigned long long *cnts)
{
    extern __shared__ double fs_lds[];
    const bool use_lds = nbins <= 2048;
    double *hs = fs_lds;
    unsigned int *hc = reinterpret_cast<unsigned int *>(fs_lds + (use_lds ? nbins : 0));
    const int s = blockIdx.y;
    if (use_lds)
        for (unsigned long i = threadIdx.x; i < nbins; i += blockDim.x) {
            hs[i] = 0.0;
            hc[i] = 0u;
        }
    __syncthreads();
    double *gs = sums + (size_t)s * nbins;
    unsigned long long *gc = cnts + (size_t)s * nbins;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < size; i += (long long)gridDim.x * blockDim.x) {
        const float2 x = a[(size_t)s * stride + i];
        float I;
        if (b) {
            const float2 y = b[(size_t)s * stride + i];
            I = sqrtf(x.x * x.x + x.y * x.y + y.x * y.x + y.y * y.y);
        } else {
            I = x.x * x.x + x.y * x.y;
        }
        if (I == 0.0f) continue;
        const unsigned long bin = fold_bin(i, nbins, t_start, dt, period);
        if (use_lds) {
            atomicAdd(&hs[bin], (double)I);
            atomicAdd(&hc[bin], 1u);
        } else {
            unsafeAtomicAdd(&gs[bin], (double)I);
            atomicAdd(&gc[bin], 1ull);
        }
    }
    __syncthreads();
    if (use_lds)
        for (unsigned long i = threadIdx.x; i < nbins; i += blockDim.x)
            if (hc[i]) {
                unsafeAtomicAdd(&gs[i], hs[i]);
                atomicAdd(&gc[i], (unsigned long long)hc[i]);
            }
}

// computeFoldedData_kernel (gpu_kernel.cu:498-508)
__global__ void profile_kernel(const double *sums, const unsigned long long *cnts, float *prof, long long n,
                               unsigned long long *cnt_out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        prof[i] = cnts[i] ? (float)(sums[i] / (double)cnts[i]) : 0.0f;
        if (cnt_out) cnt_out[i] = cnts[i];
    }
}

__global__ void fill_kernel(uint32_t *p, long long nwords, uint32_t value)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
// Workgroups per block of the split forward transform: the largest power of two that divides both
// passes' tile steps and keeps B P within the workgroups the GPU holds at once; 1 = no split.
template <int N1, int N2>
static int fwd_split_ways(int B, int cus)
{
    using S = Shape<N1, N2>;
    const int steps = (S::NTA < S::NTB ? S::NTA : S::NTB) / SUBT;
    if (S::NTA % SUBT || S::NTB % SUBT) return 1;
    int P = 1;
    while (2 * P <= steps && (long long)B * 2 * P <= cus) P *= 2;
    return P;
}

// grid workgroups of fwd_kernel, or, when every block has a scratch slot of its own (grid = B) and the
// blocks alone leave CUs idle, the two split launches over B P workgroups
template <int N1, int N2>
static hipError_t launch_fwd_t(const FwdArgs &a, int grid, int cus, hipStream_t st)
{
    using S = Shape<N1, N2>;
    const int P = grid == a.B ? fwd_split_ways<N1, N2>(a.B, cus) : 1;
    if (P == 1) {
        hipLaunchKernelGGL((fwd_kernel<N1, N2>), dim3(grid), dim3(WG), S::LDS_BYTES, st, a);
        return hipGetLastError();
    }
    const dim3 g((unsigned)(a.B * P));
    hipLaunchKernelGGL((fwd_split_kernel<N1, N2, 0>), g, dim3(WG), S::LDS_BYTES, st, a, P);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fwd_split_kernel<N1, N2, 1>), g, dim3(WG), S::LDS_BYTES, st, a, P);
    return hipGetLastError();
}

template <int N1, int N2>
static hipError_t launch_inv_t(const InvArgs &a, int grid, hipStream_t st)
{
    using S = Shape<N1, N2>;
    const size_t lds = S::LDS_BYTES + (a.fold_sum ? (size_t)a.nbins * 12 : 0) + INV_EXTRA_LDS;
    const bool ring32 = (unsigned long long)a.FBS * S::L * sizeof(float2) < (1ull << 32);
    const int var = a.tangent_ref ? VAR_TANGENT : a.out_f16 ? VAR_F16 : VAR_PLAIN;
#define LAUNCH(R, V) hipLaunchKernelGGL((inv_kernel<N1, N2, R, V>), dim3(grid), dim3(WG), lds, st, a)
    if constexpr (has_rs<N1, N2>()) {
        if ((a.order & MSOSM_SCHEDULE_SPLIT) && var == VAR_PLAIN) {
            if (ring32)
                hipLaunchKernelGGL((inv_rs_kernel<N1, true>), dim3(grid), dim3(WG), lds, st, a);
            else
                hipLaunchKernelGGL((inv_rs_kernel<N1, false>), dim3(grid), dim3(WG), lds, st, a);
            return hipGetLastError();
        }
    }
    if (ring32) {
        if (var == VAR_TANGENT) LAUNCH(true, VAR_TANGENT);
        else if (var == VAR_F16) LAUNCH(true, VAR_F16);
        else LAUNCH(true, VAR_PLAIN);
    } else {
        if (var == VAR_TANGENT) LAUNCH(false, VAR_TANGENT);
        else if (var == VAR_F16) LAUNCH(false, VAR_F16);
        else LAUNCH(false, VAR_PLAIN);
    }
#undef LAUNCH
    return hipGetLastError();
}

template <int N1, int N2>
static int lds_bytes_t()
{
    return Shape<N1, N2>::LDS_BYTES;
}


#ifndef MSOSM_SHAPES
#define MSOSM_SHAPES(X)                                                                                                \
    X(12, 64, 64) X(13, 128, 64) X(14, 128, 128) X(15, 256, 128) X(16, 256, 256) X(17, 512, 256) X(18, 512, 512)     \
    X(19, 1024, 512) X(20, 1024, 1024) X(21, 2048, 1024) X(22, 2048, 2048)
#endif

void split_length(int log2L, int *n1, int *n2)
{
    *n1 = 1 << ((log2L + 1) / 2);
    *n2 = 1 << (log2L / 2);
}

bool length_supported(int log2L) { return log2L >= 12 && log2L <= 22; }

bool split_supported(int log2L)
{
    switch (log2L) {
#define X(lg, a, b)                                                                                                    \
    case lg: return has_rs<a, b>();
        MSOSM_SHAPES(X)
#undef X
    }
    return false;
}

int tiles_lds_bytes(int log2L)
{
    switch (log2L) {
#define X(lg, a, b)                                                                                                    \
    case lg: return lds_bytes_t<a, b>();
        MSOSM_SHAPES(X)
#undef X
    }
    return -1;
}

hipError_t launch_fwd(int log2L, const FwdArgs &a, int grid, int cus, hipStream_t st)
{
    switch (log2L) {
#define X(lg, a_, b_)                                                                                                  \
    case lg: return launch_fwd_t<a_, b_>(a, grid, cus, st);
        MSOSM_SHAPES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

hipError_t launch_inv(int log2L, const InvArgs &a, int grid, hipStream_t st)
{
    if ((a.order & MSOSM_SCHEDULE_ACCUM) && accum_supported(log2L) && (a.delay_acc || a.xpre))
        return launch_accum(log2L, a, grid, st);
    switch (log2L) {
#define X(lg, a_, b_)                                                                                                  \
    case lg: return launch_inv_t<a_, b_>(a, grid, st);
        MSOSM_SHAPES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

hipError_t launch_hist(float2 *hist, const void *in, int fmt, long long first, int M, hipStream_t st)
{
    hipLaunchKernelGGL(hist_kernel, dim3((M + 255) / 256), dim3(256), 0, st, hist, in, fmt, first, M);
    return hipGetLastError();
}

hipError_t launch_bins(uint16_t *bins, long long n, unsigned long nbins, double t_start, double dt, double period,
                       hipStream_t st)
{
    hipLaunchKernelGGL(bins_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, bins, n, nbins, t_start, dt,
                       period);
    return hipGetLastError();
}

hipError_t launch_bins_sub(uint16_t *bins, long long n, unsigned long nbins, double t_start, double dt, double period,
                           int q, long long m, hipStream_t st)
{
    hipLaunchKernelGGL(bins_sub_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, bins, n, nbins, t_start,
                       dt, period, q, m);
    return hipGetLastError();
}

hipError_t launch_fold_series(const float2 *a, const float2 *b, long long size, long long stride, int nseries,
                              unsigned long nbins, double t_start, double dt, double period, double *sums,
                              unsigned long long *cnts, hipStream_t st)
{
    const size_t lds = nbins <= 2048 ? nbins * 12 : 0;
    long long blocks = (size + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fold_series_kernel, dim3((unsigned)blocks, nseries), dim3(256), lds, st, a, b, size, stride,
                       nseries, nbins, t_start, dt, period, sums, cnts);
    return hipGetLastError();
}

hipError_t launch_profile(const double *sums, const unsigned long long *cnts, float *prof, long long n, hipStream_t st,
                          unsigned long long *cnt_out)
{
    hipLaunchKernelGGL(profile_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sums, cnts, prof, n,
                       cnt_out);
    return hipGetLastError();
}

hipError_t launch_fill(void *p, long long nbytes, uint32_t value, hipStream_t st)
{
    const long long nw = nbytes / 4;
    long long blocks = (nw + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t *)p, nw, value);
    return hipGetLastError();
}

NodeDesc fwd_node(int log2L, int grid)
{
    NodeDesc n;
    switch (log2L) {
#define X(lg, a_, b_)                                                                                                  \
    case lg:                                                                                                           \
        n.func = reinterpret_cast<const void *>(&fwd_kernel<a_, b_>);                                                  \
        n.lds = Shape<a_, b_>::LDS_BYTES;                                                                              \
        break;
        MSOSM_SHAPES(X)
#undef X
    }
    n.grid = dim3(grid);
    n.block = dim3(WG);
    return n;
}

NodeDesc hist_node(int M)
{
    NodeDesc n;
    n.func = reinterpret_cast<const void *>(&hist_kernel);
    n.grid = dim3((M + 255) / 256);
    n.block = dim3(256);
    return n;
}

NodeDesc bins_node(long long cnt)
{
    NodeDesc n;
    n.func = reinterpret_cast<const void *>(&bins_kernel);
    n.grid = dim3((unsigned)((cnt + 255) / 256));
    n.block = dim3(256);
    return n;
}

NodeDesc fill_node(long long nbytes)
{
    NodeDesc n;
    long long blocks = (nbytes / 4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    n.func = reinterpret_cast<const void *>(&fill_kernel);
    n.grid = dim3((unsigned)blocks);
    n.block = dim3(256);
    return n;
}

int workgroup_threads() { return WG; }

// the role-split instances of a length take the dedispersion kernel's LDS size
template <int N1, int N2>
static hipError_t configure_rs_lds(int inv_bytes)
{
    if constexpr (has_rs<N1, N2>()) {
        for (const void *f : {reinterpret_cast<const void *>(&inv_rs_kernel<N1, true>),
                              reinterpret_cast<const void *>(&inv_rs_kernel<N1, false>)}) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, inv_bytes);
            if (e != hipSuccess) return e;
        }
    }
    (void)inv_bytes;
    return hipSuccess;
}

hipError_t configure_lds(int log2L, int fwd_bytes, int inv_bytes)
{
    switch (log2L) {
#define X(lg, a_, b_)                                                                                                  \
    case lg: {                                                                                                         \
        const void *fwd[] = {reinterpret_cast<const void *>(&fwd_kernel<a_, b_>),                                      \
                             reinterpret_cast<const void *>(&fwd_split_kernel<a_, b_, 0>),                             \
                             reinterpret_cast<const void *>(&fwd_split_kernel<a_, b_, 1>)};                            \
        hipError_t e = hipSuccess;                                                                                     \
        for (const void *f : fwd) {                                                                                    \
            e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, fwd_bytes);                         \
            if (e != hipSuccess) return e;                                                                             \
        }                                                                                                              \
        const void *inv[] = {                                                                                          \
            reinterpret_cast<const void *>(&inv_kernel<a_, b_, true, VAR_PLAIN>),                                      \
            reinterpret_cast<const void *>(&inv_kernel<a_, b_, true, VAR_F16>),                                        \
            reinterpret_cast<const void *>(&inv_kernel<a_, b_, true, VAR_TANGENT>),                                    \
            reinterpret_cast<const void *>(&inv_kernel<a_, b_, false, VAR_PLAIN>),                                     \
            reinterpret_cast<const void *>(&inv_kernel<a_, b_, false, VAR_F16>),                                       \
            reinterpret_cast<const void *>(&inv_kernel<a_, b_, false, VAR_TANGENT>)};                                  \
        for (const void *f : inv) {                                                                                    \
            e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, inv_bytes);                         \
            if (e != hipSuccess) return e;                                                                             \
        }                                                                                                              \
        return configure_rs_lds<a_, b_>(inv_bytes);                                                                    \
    }
        MSOSM_SHAPES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

} // namespace msosm
