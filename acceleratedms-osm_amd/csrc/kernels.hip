// HIP kernels of the MS-OSM coherent-dedispersion path for gfx950.
//
// Every transform of length L = N1 * N2 is a four-step FFT run by one 1024-thread workgroup:
//   pass A: N1 column sub-FFTs of length N2 (stride-N1 elements), fused with the producer (input
//           unpack, or the per-bin ring gather times the chirp H), then the W_L^{n2 k1} twiddle,
//           into the workgroup's private scratch slot;
//   pass B: N2 row sub-FFTs of length N1 from the scratch slot, fused with the consumer (ring
//           store, or discard + 1/L scale + fold / complex64 / uint16 output).
// A workgroup keeps four 4096-point tiles in flight (256 threads each). The dedispersion kernel is
// persistent: one workgroup per CU walks (DM, block) units, with the unit order chosen so that the
// workgroups sharing an XCD (and its L2) work on the same DM's chirp at the same time.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_engine.h"
#include "kernels.h"

namespace msosm {

constexpr int TILE = 4096;
constexpr int SUBT = 4;
constexpr int WG = 256 * SUBT;

template <int NS>
struct TileWords {
    // LDS float2 words one tile needs in either pass layout and in the padded sample-order buffer
    static constexpr int NC = TILE / NS;
    static constexpr int A = TileLDS<NS, NC, true>::ELEMS;
    static constexpr int B = TileLDS<NS, NC, false>::ELEMS;
    static constexpr int T = TILE + TILE / 16;
    static constexpr int AB = A > B ? A : B;
    static constexpr int V = AB > T ? AB : T;
};

template <int N1, int N2>
struct Shape {
    static constexpr int L = N1 * N2;
    static constexpr int NCA = TILE / N2 < 1 ? 1 : TILE / N2; // columns per pass-A tile
    static constexpr int NCB = TILE / N1 < 1 ? 1 : TILE / N1; // rows per pass-B tile
    static constexpr int NTA = N1 / NCA;                      // pass-A tiles
    static constexpr int NTB = N2 / NCB;                      // pass-B tiles
    static constexpr int WA = TileWords<N2>::V, WB = TileWords<N1>::V;
    static constexpr int TILE_WORDS = WA > WB ? WA : WB;
    static constexpr int LDS_TILES_BYTES = SUBT * TILE_WORDS * 8;
    // sub-FFT twiddle tables for N1 and N2 follow the tiles when they are small (<= 8 KiB; the
    // largest lengths read them from global memory); the fold histogram comes last
    static constexpr bool TW_IN_LDS = false; // measured: global (L1-resident) tables are faster here
    static constexpr int LDS_TW_BYTES = TW_IN_LDS ? (N1 + N2) * 8 : 0;
    static constexpr int LDS_BYTES = LDS_TILES_BYTES + LDS_TW_BYTES;
};



// Element idx of an array whose byte size is below 4 GiB: a 32-bit byte offset from a (uniform)
// base lets the compiler use the SGPR-base + 32-bit VGPR-offset addressing form, with no per-lane
// 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ T ld32(const T *base, uint32_t idx)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T *base, uint32_t idx, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T)) = v;
}

struct Tw {
    const float2 *n1;  // e^{2 pi i m / N1}, m < N1
    const float2 *n2;  // e^{2 pi i m / N2}, m < N2
    const float2 *lo;  // e^{2 pi i m / L}, m < 2^s
    const float2 *hi;  // e^{2 pi i m 2^s / L}, m < L / 2^s
    int s;
};

// copy the N1- and N2-point twiddle tables into LDS (read once per kernel)
template <int N1, int N2>
__device__ __forceinline__ void load_twiddles(float2 *lds_all, Tw &tw)
{
    using S = Shape<N1, N2>;
    if constexpr (!S::TW_IN_LDS) return;
    float2 *t1 = lds_all + S::LDS_TILES_BYTES / 8;
    float2 *t2 = t1 + N1;
    for (int i = threadIdx.x; i < N1; i += blockDim.x) t1[i] = tw.n1[i];
    for (int i = threadIdx.x; i < N2; i += blockDim.x) t2[i] = tw.n2[i];
    tw.n1 = t1;
    tw.n2 = t2;
    __syncthreads();
}

template <int DIR>
__device__ __forceinline__ float2 big_twiddle(const Tw &tw, int m)
{
    float2 w = cmul(tw.hi[m >> tw.s], tw.lo[m & ((1 << tw.s) - 1)]);
    return DIR < 0 ? cconj(w) : w;
}

// ---------------------------------------------------------------------------------------------
// Pass A: column sub-FFTs. load(k) returns input element k (k = k1 + N1*k2).
// Output Z[k1][n2] * W^{DIR n2 k1} goes to scr[(tile * N2 + n2) * NCA + col].
// ---------------------------------------------------------------------------------------------
template <int N1, int N2, int DIR, class Load>
__device__ __forceinline__ void pass_a(const Load &load, float2 *__restrict__ scr, float2 *lds_all, const Tw &tw)
{
    using S = Shape<N1, N2>;
    constexpr int NS = N2, NC = S::NCA, TPC = NS / 16;
#pragma unroll 1
    for (int t0 = 0; t0 < S::NTA; t0 += SUBT) {
        // opaque thread index: keeps per-thread addresses from being hoisted (and spilled) across tiles
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = tid >> 8, lt = tid & 255;
        const int col = lt % NC, jt = lt / NC;
        float2 *lds = lds_all + sub * S::TILE_WORDS;
        const int tile = t0 + sub;
        const bool active = (S::NTA % SUBT == 0) || tile < S::NTA; // all sub-tiles busy when divisible
        const int k1 = tile * NC + col;
        float2 v[16];
        if (active) {
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = load(k1 + N1 * (jt + t * TPC));
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = make_float2(0.f, 0.f);
        }
        subfft<NS, NC, true, DIR>(v, lds, col, jt, tw.n2);
        if (active) {
            // element (q, t) is output n2 = jt + q*TPC + t*NsL; its twiddle W^{n2 k1} is
            // W^{(jt + q*TPC) k1} (table) times c^t with c = W^{NsL k1} from binary powers, so a
            // thread does 2 + 2*NB table reads per tile instead of two per element
            constexpr int R = FinalShape<NS>::R, NsL = NS / R;
            const float2 c = big_twiddle<DIR>(tw, k1 * NsL);
#pragma unroll
            for (int q = 0; q < 16 / R; ++q) {
                float2 w = big_twiddle<DIR>(tw, (jt + q * TPC) * k1);
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const int n2 = final_pos<NS>(jt, q, t);
                    st32(scr, (uint32_t)((tile * N2 + n2) * NC + col), cmul(v[q * R + t], w));
                    if (t + 1 < R) w = cmul(w, c); // depth <= 15 products: < 1e-6 relative
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Pass B: row sub-FFTs of the scratch slot. For each tile, emit(lds, tile, active) is called with
// the tile's outputs in LDS in padded sample order: lds[lds_pad(f)] = y[n2 + N2 * n1] for
// f = n1 * NCB + row, n2 = tile * NCB + row. Epilogues read f = lane + 256 * i, so consecutive
// lanes hold consecutive samples (runs of NCB) and neither the LDS reads nor the stores conflict.
// ---------------------------------------------------------------------------------------------
template <int N1, int N2, int DIR, class Emit>
__device__ __forceinline__ void pass_b(const float2 *__restrict__ scr, const Emit &emit, float2 *lds_all, const Tw &tw)
{
    using S = Shape<N1, N2>;
    constexpr int NS = N1, NC = S::NCB, TPC = NS / 16, NCA = S::NCA;
#pragma unroll 1
    for (int t0 = 0; t0 < S::NTB; t0 += SUBT) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = tid >> 8, lt = tid & 255;
        const int jt = lt % TPC, row = lt / TPC;
        float2 *lds = lds_all + sub * S::TILE_WORDS;
        const int tile = t0 + sub;
        const bool active = (S::NTB % SUBT == 0) || tile < S::NTB;
        const int n2 = tile * NC + row;
        float2 v[16];
        if (active) {
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int k1 = jt + t * TPC;
                v[t] = ld32(scr, (uint32_t)(((k1 / NCA) * N2 + n2) * NCA + (k1 % NCA)));
            }
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = make_float2(0.f, 0.f);
        }
        subfft<NS, NC, false, DIR>(v, lds, row, jt, tw.n1);
        __syncthreads();
        {
            constexpr int R = FinalShape<NS>::R;
#pragma unroll
            for (int q = 0; q < 16 / R; ++q)
#pragma unroll
                for (int t = 0; t < R; ++t) lds[lds_pad(final_pos<NS>(jt, q, t) * NC + row)] = v[q * R + t];
        }
        __syncthreads();
        emit(lds, tile, active);
    }
}

// ---------------------------------------------------------------------------------------------
// Input decoding
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 decode_in(const void *in, int fmt, long long s)
{
    // one push is far below 4 GiB, so the sample offset is a 32-bit byte offset
    if (fmt == MSOSM_IN_U16) {
        // uint16 pair, offset binary with code 0 reserved for 0 (gpu_kernel.cu:22-29)
        const uint32_t u = ld32(reinterpret_cast<const uint32_t *>(in), (uint32_t)s);
        const uint32_t a = u & 0xFFFFu, b = u >> 16;
        return make_float2((float)a - (a ? 32768.f : 0.f), (float)b - (b ? 32768.f : 0.f));
    } else if (fmt == MSOSM_IN_I8) {
        const char2 c = ld32(reinterpret_cast<const char2 *>(in), (uint32_t)s);
        return make_float2((float)c.x, (float)c.y);
    }
    return ld32(reinterpret_cast<const float2 *>(in), (uint32_t)s);
}

// ---------------------------------------------------------------------------------------------
// Forward: unpack + window + DFT_L into ring slots. One unit = one block of M new samples.
// ---------------------------------------------------------------------------------------------
template <int N1, int N2>
__global__ void __launch_bounds__(WG) fwd_kernel(FwdArgs a)
{
    using S = Shape<N1, N2>;
    constexpr int L = S::L, M = L / 2;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N1, N2>(lds_all, tw);
    float2 *scr = a.scratch + (size_t)blockIdx.x * L;
#pragma unroll 1
    for (int b = blockIdx.x; b < a.B; b += gridDim.x) {
        const long long base = (long long)(b - 1) * M;
        auto load = [&](int k) -> float2 {
            const long long s = base + k;
            if (s < 0) return a.hist[s + M];
            return decode_in(a.in, a.in_fmt, s);
        };
        pass_a<N1, N2, -1>(load, scr, lds_all, tw);
        __syncthreads();
        const long long slot = (a.g0 + b) % a.FBS;
        float2 *dst = a.ring + (size_t)slot * L;
        auto emit = [&](const float2 *lds, int tile, bool active) {
            if (!active) return;
            constexpr int NC = S::NCB;
            const int lt = threadIdx.x & 255;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int f = lt + 256 * i;
                const int n1 = f / NC, row = f % NC;
                st32(dst, (uint32_t)(tile * NC + row + N2 * n1), lds[lds_pad(f)]);
            }
        };
        pass_b<N1, N2, -1>(scr, emit, lds_all, tw);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// Dedispersion: gather * H -> IDFT_L -> keep [M, L) * 1/L -> fold / complex64 / uint16.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t enc_u16(float x)
{
    // round-to-nearest-even of x + 32768*[x != 0], clamped to [0, 65535] (NaN -> 0)
    float v = x + (x != 0.f ? 32768.f : 0.f);
    v = fminf(fmaxf(v, 0.f), 65535.f);
    return (uint32_t)rintf(v);
}

// FP16 pair of y / scale (exact power-of-two scaling, round-to-nearest-even, saturated at +-65504)
__device__ __forceinline__ uint32_t enc_f16_pair(float2 y, float inv_scale)
{
    const float xr = fminf(fmaxf(y.x * inv_scale, -65504.f), 65504.f);
    const float xi = fminf(fmaxf(y.y * inv_scale, -65504.f), 65504.f);
    const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
    return (uint32_t)__builtin_bit_cast(uint16_t, hr) | ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
}

template <int N1, int N2, bool RING32>
__global__ void __launch_bounds__(WG) inv_kernel(InvArgs a)
{
    using S = Shape<N1, N2>;
    constexpr int L = S::L, M = L / 2;
    extern __shared__ float2 lds_all[];
    Tw tw{a.tw_n1, a.tw_n2, a.tw_lo, a.tw_hi, a.tw_s};
    load_twiddles<N1, N2>(lds_all, tw);
    float2 *scr = a.scratch + (size_t)blockIdx.x * L;
    const bool fold = a.fold_sum != nullptr;
    const bool f16 = a.out_f16 != nullptr;
    double *hsum = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_all) + S::LDS_BYTES);
    unsigned int *hcnt = reinterpret_cast<unsigned int *>(hsum + a.nbins);
    float *pw_lds = reinterpret_cast<float *>(hcnt + a.nbins + (a.nbins & 1)); // FP16 scale reduction
    if (fold) {
        for (int i = threadIdx.x; i < a.nbins; i += WG) {
            hsum[i] = 0.0;
            hcnt[i] = 0u;
        }
    }
    // XCD-aware virtual id: workgroups w, w+8, ... share an XCD; give them consecutive virtual ids
    const int G = gridDim.x;
    const int w = blockIdx.x;
    const int v = (G % 8 == 0) ? ((w % 8) * (G / 8) + w / 8) : w;
    const long long U = (long long)a.D * a.B;
    const float scale = 1.0f / (float)L;
    if (f16 && threadIdx.x == 0) *pw_lds = 0.f;
#pragma unroll 1
    for (long long u = v; u < U; u += G) {
        const int d = (int)(u / a.B), b = (int)(u % a.B);
        const int gm = (int)((a.g0 + b) % a.FBS); // ring slot of this block
        const int FBS = a.FBS;
        const float2 *__restrict__ Hd = a.H + (size_t)d * L;
        const uint16_t *__restrict__ dl = a.delay + (size_t)d * L;
        const float2 *__restrict__ ring = a.ring;
        float power = 0.f; // sum |Y|^2 of this thread's bins (FP16 output scale, Parseval)
        auto load = [&](int k) -> float2 {
            int src = gm - (int)ld32(dl, (uint32_t)k); // delays are < FBS, so one wrap at most
            if (src < 0) src += FBS;
            const float2 x = RING32 ? ld32(ring, (uint32_t)src * (uint32_t)L + (uint32_t)k) : ring[(size_t)src * L + k];
            const float2 y = cmul(x, ld32(Hd, (uint32_t)k));
            if (f16) power += y.x * y.x + y.y * y.y;
            return y;
        };
        pass_a<N1, N2, +1>(load, scr, lds_all, tw);
        float inv_scale = 1.f, out_scale = 1.f;
        if (f16) {
            // block rms of the 1/L-scaled series: sqrt(sum |Y|^2) / L; scale = 2^floor(log2(rms))
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) power += __shfl_xor(power, off, 64);
            if ((threadIdx.x & 63) == 0) atomicAdd(pw_lds, power);
            __syncthreads();
            const float tot = *pw_lds;
            const float rms = sqrtf(tot) * scale;
            int e2 = 0;
            (void)frexpf(rms, &e2); // rms = m * 2^e2, m in [0.5, 1)
            out_scale = rms > 0.f ? ldexpf(1.f, e2 - 1) : 1.f;
            inv_scale = 1.f / out_scale; // exact: a power of two
            if (threadIdx.x == 0 && a.out_scale) a.out_scale[(size_t)d * a.B + b] = out_scale;
        }
        __syncthreads();
        float2 *oc = a.out_c64 ? a.out_c64 + ((size_t)d * a.B + b) * M : nullptr;
        uint32_t *ou = a.out_u16 ? a.out_u16 + ((size_t)d * a.B + b) * M : nullptr;
        uint32_t *oh = a.out_f16 ? a.out_f16 + ((size_t)d * a.B + b) * M : nullptr;
        const uint16_t *bins = a.bins ? a.bins + (size_t)b * M : nullptr;
        auto emit = [&](const float2 *lds, int tile, bool active) {
            if (!active) return;
            constexpr int NC = S::NCB;
            const int lt = threadIdx.x & 255;
            // upper half of the tile's outputs (n1 >= N1/2): 8 consecutive f per thread, i.e. runs
            // of consecutive samples; the fold accumulates a run in registers and flushes on a bin change
            const int fbase = (N1 / 2) * NC + lt * 8;
            int cur = -1;
            double csum = 0.0;
            unsigned int ccnt = 0;
            // runs of >= 8 consecutive samples: their 8 bins are one aligned 16-byte load
            uint16_t bv[8];
            if (fold && NC >= 8) {
                const int n1 = fbase / NC, row = fbase % NC;
                const uint4 b4 = *reinterpret_cast<const uint4 *>(bins + (tile * NC + row + N2 * n1 - M));
                const uint32_t w4[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    bv[2 * i] = (uint16_t)(w4[i] & 0xFFFFu);
                    bv[2 * i + 1] = (uint16_t)(w4[i] >> 16);
                }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int f = fbase + i;
                const int n1 = f / NC, row = f % NC;
                const int m = tile * NC + row + N2 * n1 - M;
                float2 y = lds[lds_pad(f)];
                y.x *= scale;
                y.y *= scale;
                if (oc) st32(oc, (uint32_t)m, y);
                if (ou) st32(ou, (uint32_t)m, (enc_u16(y.y) << 16) | enc_u16(y.x));
                if (oh) st32(oh, (uint32_t)m, enc_f16_pair(y, inv_scale));
                if (fold) {
                    const float I = y.x * y.x + y.y * y.y;
                    if (I != 0.0f) {
                        const int bin = NC >= 8 ? (int)bv[i] : (int)bins[m];
                        if (bin != cur) {
                            if (ccnt) {
                                atomicAdd(&hsum[cur], csum);
                                atomicAdd(&hcnt[cur], ccnt);
                            }
                            cur = bin;
                            csum = 0.0;
                            ccnt = 0;
                        }
                        csum += (double)I;
                        ccnt += 1;
                    }
                }
            }
            if (fold && ccnt) {
                atomicAdd(&hsum[cur], csum);
                atomicAdd(&hcnt[cur], ccnt);
            }
        };
        pass_b<N1, N2, +1>(scr, emit, lds_all, tw);
        __syncthreads();
        if (f16 && threadIdx.x == 0) *pw_lds = 0.f; // next unit's reduction (ordered by the barriers)
        if (fold) {
            double *gs = a.fold_sum + (size_t)d * a.nbins;
            unsigned long long *gc = a.fold_cnt + (size_t)d * a.nbins;
            for (int i = threadIdx.x; i < a.nbins; i += WG) {
                const unsigned int c = hcnt[i];
                if (c) {
                    unsafeAtomicAdd(&gs[i], hsum[i]);
                    atomicAdd(&gc[i], (unsigned long long)c);
                    hsum[i] = 0.0;
                    hcnt[i] = 0u;
                }
            }
            __syncthreads();
        }
    }
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
__global__ void profile_kernel(const double *sums, const unsigned long long *cnts, float *prof, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prof[i] = cnts[i] ? (float)(sums[i] / (double)cnts[i]) : 0.0f;
}

__global__ void fill_kernel(uint32_t *p, long long nwords, uint32_t value)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (long long)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
template <int N1, int N2>
static hipError_t launch_fwd_t(const FwdArgs &a, int grid, hipStream_t st)
{
    using S = Shape<N1, N2>;
    hipLaunchKernelGGL((fwd_kernel<N1, N2>), dim3(grid), dim3(WG), S::LDS_BYTES, st, a);
    return hipGetLastError();
}

template <int N1, int N2>
static hipError_t launch_inv_t(const InvArgs &a, int grid, hipStream_t st)
{
    using S = Shape<N1, N2>;
    const size_t lds = S::LDS_BYTES + (a.fold_sum ? (size_t)a.nbins * 12 : 0) + INV_EXTRA_LDS;
    const bool ring32 = (unsigned long long)a.FBS * S::L * sizeof(float2) < (1ull << 32);
    if (ring32)
        hipLaunchKernelGGL((inv_kernel<N1, N2, true>), dim3(grid), dim3(WG), lds, st, a);
    else
        hipLaunchKernelGGL((inv_kernel<N1, N2, false>), dim3(grid), dim3(WG), lds, st, a);
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

hipError_t launch_fwd(int log2L, const FwdArgs &a, int grid, hipStream_t st)
{
    switch (log2L) {
#define X(lg, a_, b_)                                                                                                  \
    case lg: return launch_fwd_t<a_, b_>(a, grid, st);
        MSOSM_SHAPES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

hipError_t launch_inv(int log2L, const InvArgs &a, int grid, hipStream_t st)
{
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

hipError_t launch_profile(const double *sums, const unsigned long long *cnts, float *prof, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(profile_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sums, cnts, prof, n);
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

int workgroup_threads() { return WG; }

hipError_t configure_lds(int log2L, int fwd_bytes, int inv_bytes)
{
    switch (log2L) {
#define X(lg, a_, b_)                                                                                                  \
    case lg: {                                                                                                         \
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&fwd_kernel<a_, b_>),                        \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, fwd_bytes);                     \
        if (e != hipSuccess) return e;                                                                                 \
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(&inv_kernel<a_, b_, true>),                             \
                                hipFuncAttributeMaxDynamicSharedMemorySize, inv_bytes);                                \
        if (e != hipSuccess) return e;                                                                                 \
        return hipFuncSetAttribute(reinterpret_cast<const void *>(&inv_kernel<a_, b_, false>),                         \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, inv_bytes);                             \
    }
        MSOSM_SHAPES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

} // namespace msosm
