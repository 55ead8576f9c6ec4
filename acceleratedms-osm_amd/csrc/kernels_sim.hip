// Batched FFTs of the GPU pulsar simulator (dispersion injection, simulated_complex_signal.cpp:74-277).
// One launch transforms `ntrans` sequences of one length L = N1 * N2 with the same in-workgroup
// four-step engine as the dedispersion kernels: transform t reads its element e at
// src[t * in_t + e * in_e] and sends its output o through one fused output stage (plain store, the
// twiddle of a six-step decomposition, the chirp multiply, or scale + extract + uint16 encode).
#include "msosm_device.h"

namespace msosm {

#define MSOSM_SIM_SHAPES(X)                                                                                            \
    X(12, 64, 64) X(13, 128, 64) X(14, 128, 128) X(15, 256, 128) X(16, 256, 256) X(17, 512, 256) X(18, 512, 512)     \
    X(19, 1024, 512) X(20, 1024, 1024) X(21, 2048, 1024) X(22, 2048, 2048)

// x + 32768 converted as the reference's static_cast<uint16_t>(float) does on x86-64 (truncation to
// a 32-bit integer, low 16 bits kept)
__device__ __forceinline__ uint32_t sim_u16(float x) { return (uint32_t)(int)(x + 32768.0f) & 0xFFFFu; }

template <int N1, int N2, int DIR>
__global__ void __launch_bounds__(WG) sim_fft_kernel(SimFftArgs a)
{
    using S = Shape<N1, N2>;
    constexpr int L = S::L;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N1, N2>(lds_all, tw);
    float2 *scr = a.scratch + (size_t)blockIdx.x * L;
#pragma unroll 1
    for (long long t = blockIdx.x; t < a.ntrans; t += gridDim.x) {
        const float2 *src = a.src + t * a.in_t;
        auto load = [&](int k) -> float2 { return src[(long long)k * a.in_e]; };
        pass_a<N1, N2, DIR>(load, scr, lds_all, tw);
        __syncthreads();
        auto prep = [](int, bool) { return 0; };
        auto emit = [&](const float2 *lds, int tile, bool active, int) {
            if (!active) return;
            constexpr int NC = S::NCB;
            const int lt = threadIdx.x & 255;
#pragma unroll 4
            for (int i = 0; i < 16; ++i) {
                const int f = lt + 256 * i;
                const int o = tile * NC + f % NC + N2 * (f / NC); // output index
                float2 y = lds[lds_pad(lt) + 272 * i];           // pad(lt + 256 i)
                const long long di = t * a.out_t + (long long)o * a.out_e;
                if (a.mode == SIM_TWIDDLE) {
                    // e^{DIR 2 pi i (o t mod N) / N} in FP64 (exact argument: the ratio is a dyadic)
                    const long long m = ((long long)o * t) & (a.big_n - 1);
                    double s, c;
                    sincospi(2.0 * (double)m / (double)a.big_n, &s, &c);
                    y = cmul(y, make_float2((float)c, DIR < 0 ? -(float)s : (float)s));
                    a.dst[di] = y;
                } else if (a.mode == SIM_CHIRP) {
                    a.dst[di] = cmul(y, a.chirp[di]);
                } else if (a.mode == SIM_EXTRACT) {
                    const long long j = di - a.ex_start;
                    if (j >= 0 && j < a.ex_count) {
                        // the unnormalised inverse transform, then times 1/N as two FP32 products
                        const float2 v = make_float2(y.x * a.ex_scale, y.y * a.ex_scale);
                        if (a.ex_c64) a.ex_c64[j] = v;
                        if (a.ex_u16) a.ex_u16[j] = sim_u16(v.x) | (sim_u16(v.y) << 16);
                    }
                } else {
                    a.dst[di] = y;
                }
            }
        };
        pass_b<N1, N2, DIR>(scr, prep, emit, lds_all, tw);
        __syncthreads();
    }
}

bool sim_length_supported(int log2L) { return log2L >= 12 && log2L <= 22; }

hipError_t launch_sim_fft(int log2L, int dir, const SimFftArgs &a, int grid, hipStream_t st)
{
    switch (log2L) {
#define X(lg, n1, n2)                                                                                                  \
    case lg: {                                                                                                         \
        const int lds = Shape<n1, n2>::LDS_BYTES;                                                                      \
        const void *f = dir < 0 ? reinterpret_cast<const void *>(&sim_fft_kernel<n1, n2, -1>)                          \
                                : reinterpret_cast<const void *>(&sim_fft_kernel<n1, n2, +1>);                         \
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);                        \
        if (e != hipSuccess) return e;                                                                                 \
        if (dir < 0)                                                                                                   \
            hipLaunchKernelGGL((sim_fft_kernel<n1, n2, -1>), dim3(grid), dim3(WG), lds, st, a);                        \
        else                                                                                                           \
            hipLaunchKernelGGL((sim_fft_kernel<n1, n2, +1>), dim3(grid), dim3(WG), lds, st, a);                        \
        return hipGetLastError();                                                                                      \
    }
        MSOSM_SIM_SHAPES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

} // namespace msosm
