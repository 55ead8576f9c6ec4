// Device building blocks of the MS-OSM kernels for gfx950: the four-step FFT passes of one
// 1024-thread workgroup, input decoding, output codecs and the dedispersion unit.
//
// Every transform of length L = N1 * N2 is a four-step FFT:
//   pass A: N1 column sub-FFTs of length N2 (stride-N1 elements), fused with the producer (input
//           unpack, or the per-bin ring gather times the chirp H), then the W_L^{n2 k1} twiddle,
//           into the workgroup's private scratch slot;
//   pass B: N2 row sub-FFTs of length N1 from the scratch slot, fused with the consumer (ring
//           store, or discard + fold / complex64 / uint16 / FP16 output; 1/L rides in H).
// A workgroup keeps four 4096-point tiles in flight (256 threads each).
#pragma once
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

template <int N1, int N2, int NSUB = SUBT>
struct Shape {
    static constexpr int L = N1 * N2;
    static constexpr int NCA = TILE / N2 < 1 ? 1 : TILE / N2; // columns per pass-A tile
    static constexpr int NCB = TILE / N1 < 1 ? 1 : TILE / N1; // rows per pass-B tile
    static constexpr int NTA = N1 / NCA;                      // pass-A tiles
    static constexpr int NTB = N2 / NCB;                      // pass-B tiles
    static constexpr int WA = TileWords<N2>::V, WB = TileWords<N1>::V;
    static constexpr int TILE_WORDS = WA > WB ? WA : WB;
    static constexpr int LDS_TILES_BYTES = NSUB * TILE_WORDS * 8;
    // The twiddle tables (sub-FFT N1 and N2, four-step lo and hi: 2 (N1 + N2) entries) follow the
    // tiles in LDS where they fit beside a 1024-bin fold histogram: their reads then count on the LDS
    // counter and never wait for the prefetched next tile's global loads. Otherwise they stay in
    // global memory (L1-resident). The fold histogram comes last.
    static constexpr int TW_BYTES = 2 * (N1 + N2) * 8;
    static constexpr bool TW_IN_LDS = NSUB * TILE_WORDS * 8 + TW_BYTES + 1024 * 12 + 16 <= 160 * 1024;
    static constexpr int LDS_TW_BYTES = TW_IN_LDS ? TW_BYTES : 0;
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

// copy the twiddle tables into LDS (read once per kernel) where the shape keeps them there. The
// four-step tables are the split lo = e^{2 pi i m/L} (m < 2^s), hi = e^{2 pi i m 2^s/L} with
// 2^s = N1, the split the host builds for every length (checked at plan creation). All four
// pointers then refer to LDS only, so their reads compile to LDS instructions (a pointer that may
// refer to either memory compiles to flat loads, which wait on both memory counters).
template <int N1, int N2, int NSUB = SUBT>
__device__ __forceinline__ void load_twiddles(float2 *lds_all, Tw &tw)
{
    using S = Shape<N1, N2, NSUB>;
    if constexpr (S::TW_IN_LDS) {
        float2 *t1 = lds_all + S::LDS_TILES_BYTES / 8;
        float2 *t2 = t1 + N1, *tl = t2 + N2, *th = tl + N1;
        for (int i = threadIdx.x; i < N1; i += blockDim.x) {
            t1[i] = tw.n1[i];
            tl[i] = tw.lo[i];
        }
        for (int i = threadIdx.x; i < N2; i += blockDim.x) {
            t2[i] = tw.n2[i];
            th[i] = tw.hi[i];
        }
        tw.n1 = t1;
        tw.n2 = t2;
        tw.lo = tl;
        tw.hi = th;
        __syncthreads();
    }
}

template <int DIR>
__device__ __forceinline__ float2 big_twiddle(const Tw &tw, int m)
{
    float2 w = cmul(tw.hi[m >> tw.s], tw.lo[m & ((1 << tw.s) - 1)]);
    return DIR < 0 ? cconj(w) : w;
}

// Position of element (k1, n2) of the four-step intermediate in a workgroup's scratch slot (pass A
// columns k1 < N1, rows n2 < N2; NC columns per pass-A tile, NC rows per pass-B tile).
//   layout 0: [k1 / NC][n2 / NC][n2 % NC][k1 % NC]
//   layout 1: [n2 / NC][k1 / NC][n2 % NC][k1 % NC]   (square lengths, N1 = N2)
// Pass B's tile t (rows n2 / NC = t) of layout 0 occupies exactly the positions of pass A's tile t
// (columns k1 / NC = t) of layout 1, and the other way round: a pass A writing one layout can take
// over, tile by tile, the positions that a pass B reading the other layout has consumed. Both keep
// runs of NC consecutive k1 (the lanes of a store or load) contiguous.
template <int V>
struct IntC {
    static constexpr int value = V;
};

template <int N2, int NC>
__device__ __forceinline__ uint32_t slot_pos(int k1, int n2, int layout)
{
    if (layout == 0) return (uint32_t)(((k1 / NC) * N2 + n2) * NC + (k1 % NC));
    return (uint32_t)((n2 / NC) * (N2 * NC) + (k1 / NC) * (NC * NC) + (n2 % NC) * NC + (k1 % NC));
}

// ---------------------------------------------------------------------------------------------
// Pass A: column sub-FFTs. load(k) returns input element k (k = k1 + N1*k2).
// Output Z[k1][n2] * W^{DIR n2 k1} goes to scr[(tile * N2 + n2) * NCA + col].
// ---------------------------------------------------------------------------------------------
// Materialises v here: keeps the compiler from sinking the computation of v below the next tile's
// (conditional) loads, which would keep the consumed registers alive across them (and copy them).
__device__ __forceinline__ void pin16(float2 (&v)[16])
{
#pragma unroll
    for (int t = 0; t < 16; ++t) asm volatile("" : "+v"(v[t].x), "+v"(v[t].y));
}

// One pass-A tile step: the column sub-FFTs of the tile this thread's 256-thread group works on,
// then the W^{n2 k1} twiddle and the store into the scratch slot (inactive groups only take part in
// the barriers).
template <int N1, int N2, int DIR, int NSUB>
__device__ __forceinline__ void pass_a_tile(float2 (&v)[16], float2 *__restrict__ scr, float2 *lds_all, const Tw &tw,
                                            int t0, int layout = 0)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int NS = N2, NC = S::NCA, TPC = NS / 16;
    // opaque thread index: keeps per-thread addresses from being hoisted (and spilled) across tiles
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int sub = tid >> 8, lt = tid & 255;
    const int col = lt % NC, jt = lt / NC;
    float2 *lds = lds_all + sub * S::TILE_WORDS;
    const int tile = t0 + sub;
    const bool active = (S::NTA % NSUB == 0) || tile < S::NTA;
    const int k1 = tile * NC + col;
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
                st32(scr, slot_pos<N2, NC>(k1, n2, layout), cmul(v[q * R + t], w));
                if (t + 1 < R) w = cmul(w, c); // depth <= 15 products: < 1e-6 relative
            }
        }
    }
}

// With PF, the next tile's inputs are loaded while the current tile is transformed (32 more VGPRs):
// the loads' latency then overlaps the sub-FFT instead of stalling every wave at the tile start.
// Tiles [t_begin, t_end) (multiples of NSUB; by default all of them).
template <int N1, int N2, int DIR, bool PF = false, int NSUB = SUBT, class Load>
__device__ __forceinline__ void pass_a(const Load &load, float2 *__restrict__ scr, float2 *lds_all, const Tw &tw,
                                       int t_begin = 0, int t_end = Shape<N1, N2, NSUB>::NTA)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int NS = N2, NC = S::NCA, TPC = NS / 16;
    auto fetch = [&](int t0, float2 *dst) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = tid >> 8, lt = tid & 255;
        const int col = lt % NC, jt = lt / NC;
        const int tile = t0 + sub;
        const bool active = (S::NTA % NSUB == 0) || tile < S::NTA; // all sub-tiles busy when divisible
        const int k1 = tile * NC + col;
        if (active) {
#pragma unroll
            for (int t = 0; t < 16; ++t) dst[t] = load(k1 + N1 * (jt + t * TPC));
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) dst[t] = make_float2(0.f, 0.f);
        }
    };
    float2 nx[16];
    if constexpr (PF) fetch(t_begin, nx);
#pragma unroll 1
    for (int t0 = t_begin; t0 < t_end; t0 += NSUB) {
        float2 v[16];
        if constexpr (PF) {
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = nx[t];
            if (t0 + NSUB < t_end) fetch(t0 + NSUB, nx);
        } else {
            fetch(t0, v);
        }
        pass_a_tile<N1, N2, DIR, NSUB>(v, scr, lds_all, tw, t0);
    }
}

// Pass A of the dedispersion unit with the per-bin delays of the next tile in flight (16 VGPRs): at
// each tile step the ring and chirp loads are issued at once, from delays already in registers, so a
// step waits for one memory round trip instead of two dependent ones.
// XB extra barriers follow each step (the role-split kernel matches the pass-B role's count).
template <int N1, int N2, bool RING32, bool POWER, int NSUB, int XB = 0>
__device__ __forceinline__ void pass_a_gather_dl(const float2 *__restrict__ ring, int FBS, int gm,
                                                 const float2 *__restrict__ Hd, const uint16_t *__restrict__ dl,
                                                 float2 *__restrict__ scr, float2 *lds_all, const Tw &tw, float &power,
                                                 int layout = 0)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int L = S::L, NC = S::NCA, TPC = N2 / 16, NT = S::NTA;
    constexpr bool FULL = NT % NSUB == 0;
    constexpr int LAST = (NT - 1) / NSUB * NSUB;
    int dd[16];
    auto load_delays = [&](int t0) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = tid >> 8, lt = tid & 255;
        const int tile = t0 + sub;
        if (FULL || tile < NT) {
            const uint32_t k0 = (uint32_t)(lt % NC + N1 * (lt / NC) + tile * NC);
#pragma unroll
            for (int t = 0; t < 16; ++t) dd[t] = (int)ld32(dl, k0 + (uint32_t)(N1 * TPC * t));
        }
    };
    load_delays(0);
#pragma unroll 1
    for (int t0 = 0; t0 < NT; t0 += NSUB) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = tid >> 8, lt = tid & 255;
        const int tile = t0 + sub;
        float2 x[16], h[16];
        if (FULL || tile < NT) {
            const uint32_t k0 = (uint32_t)(lt % NC + N1 * (lt / NC) + tile * NC);
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const uint32_t k = k0 + (uint32_t)(N1 * TPC * t);
                int src = gm - dd[t]; // delays are < FBS, so one wrap at most
                if (src < 0) src += FBS;
                x[t] = RING32 ? ld32(ring, (uint32_t)src * (uint32_t)L + k) : ring[(size_t)src * L + k];
                h[t] = ld32(Hd, k);
            }
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) x[t] = h[t] = make_float2(0.f, 0.f);
        }
        load_delays(min(t0 + NSUB, LAST)); // the last step re-reads its own delays (L2)
        float2 v[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            v[t] = cmul(x[t], h[t]);
            if constexpr (POWER) power += v[t].x * v[t].x + v[t].y * v[t].y;
        }
        pass_a_tile<N1, N2, +1, NSUB>(v, scr, lds_all, tw, t0, layout);
#pragma unroll
        for (int i = 0; i < XB; ++i) __syncthreads();
    }
}

// Pass A of the role-split kernel: the ring values of the next tile step are loaded right after this
// step's slot stores, so they are in flight across the step's last barriers (where the pass-B role
// writes and emits its tile) instead of being waited for at the start of the next step; the chirp
// values (mostly L2 hits) are loaded at the step start with the delays of the step after. The 32
// VGPRs of the early ring values are free at that point (the transform's values are stored). The
// last step loads the first delays and ring values of the workgroup's next unit (next: there is
// one, at ring slot gm_next with delay table dl_next), so x and dd carry over from one call to the
// next and only the first unit (first) loads its own. XB extra barriers follow each step.
template <int N1, int N2, bool RING32, int NSUB, int XB>
__device__ __forceinline__ void pass_a_gather_pf(const float2 *__restrict__ ring, int FBS, int gm,
                                                 const float2 *__restrict__ Hd, const uint16_t *__restrict__ dl,
                                                 float2 *__restrict__ scr, float2 *lds_all, const Tw &tw, int layout,
                                                 float2 (&x)[16], int (&dd)[16], bool first, bool next, int gm_next,
                                                 const uint16_t *__restrict__ dl_next)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int L = S::L, NC = S::NCA, TPC = N2 / 16, NT = S::NTA;
    static_assert(NT % NSUB == 0, "whole tile steps");
    auto first_bin = [&](int t0) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lt = tid & 255, tile = t0 + (tid >> 8);
        return (uint32_t)(lt % NC + N1 * (lt / NC) + tile * NC);
    };
    auto load_delays = [&](const uint16_t *__restrict__ tab, int t0) {
        const uint32_t k0 = first_bin(t0);
#pragma unroll
        for (int t = 0; t < 16; ++t) dd[t] = (int)ld32(tab, k0 + (uint32_t)(N1 * TPC * t));
    };
    auto gather = [&](int g, int t0) {
        const uint32_t k0 = first_bin(t0);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const uint32_t k = k0 + (uint32_t)(N1 * TPC * t);
            int src = g - dd[t]; // delays are < FBS, so one wrap at most
            if (src < 0) src += FBS;
            x[t] = RING32 ? ld32(ring, (uint32_t)src * (uint32_t)L + k) : ring[(size_t)src * L + k];
        }
    };
    if (first) {
        load_delays(dl, 0);
        gather(gm, 0);
    }
#pragma unroll 1
    for (int t0 = 0; t0 < NT; t0 += NSUB) {
        float2 h[16];
        {
            const uint32_t k0 = first_bin(t0);
#pragma unroll
            for (int t = 0; t < 16; ++t) h[t] = ld32(Hd, k0 + (uint32_t)(N1 * TPC * t));
        }
        const bool last = t0 + NSUB >= NT;
        if (!last)
            load_delays(dl, t0 + NSUB);
        else if (next)
            load_delays(dl_next, 0);
        float2 v[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) v[t] = cmul(x[t], h[t]);
        pass_a_tile<N1, N2, +1, NSUB>(v, scr, lds_all, tw, t0, layout);
        // compiler-only memory barrier: the next loads may not move above the stores (with the
        // restrict pointers they could, and would then be live across the transform)
        asm volatile("" ::: "memory");
        if (!last)
            gather(gm, t0 + NSUB);
        else if (next)
            gather(gm_next, 0);
#pragma unroll
        for (int i = 0; i < XB; ++i) __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// Pass B: row sub-FFTs of the scratch slot. For each tile, emit(lds, tile, active) is called with
// the tile's outputs in LDS in padded sample order: lds[lds_pad(f)] = y[n2 + N2 * n1] for
// f = n1 * NCB + row, n2 = tile * NCB + row. Epilogues read f = lane + 256 * i, so consecutive
// lanes hold consecutive samples (runs of NCB) and neither the LDS reads nor the stores conflict.
// ---------------------------------------------------------------------------------------------
// prep(tile, active) runs at the start of each tile step, before the next tile's prefetch is issued
// (global loads the epilogue needs: they retire before the prefetch, so waiting for them does not
// wait for it); its result is handed to emit(lds, tile, active, prepared).
// SUB0: the first of this pass's 256-thread groups (the role-split kernel runs pass B on groups 2, 3).
// Tiles [t_begin, t_end) (multiples of NSUB; by default all of them).
template <int N1, int N2, int DIR, bool PF = false, int NSUB = SUBT, int SUB0 = 0, class Prep, class Emit>
__device__ __forceinline__ void pass_b(const float2 *__restrict__ scr, const Prep &prep, const Emit &emit,
                                       float2 *lds_all, const Tw &tw, int layout = 0, int t_begin = 0,
                                       int t_end = Shape<N1, N2, NSUB>::NTB)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int NS = N1, NC = S::NCB, TPC = NS / 16, NCA = S::NCA;
    auto fetch = [&](int t0, float2 *dst) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = (tid >> 8) - SUB0, lt = tid & 255;
        const int jt = lt % TPC, row = lt / TPC;
        const int tile = t0 + sub;
        const bool active = (S::NTB % NSUB == 0) || tile < S::NTB;
        const int n2 = tile * NC + row;
        if (active) {
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int k1 = jt + t * TPC;
                dst[t] = ld32(scr, slot_pos<N2, NCA>(k1, n2, layout));
            }
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) dst[t] = make_float2(0.f, 0.f);
        }
    };
    float2 nx[16];
    if constexpr (PF) fetch(t_begin, nx);
#pragma unroll 1
    for (int t0 = t_begin; t0 < t_end; t0 += NSUB) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = (tid >> 8) - SUB0, lt = tid & 255;
        const int jt = lt % TPC, row = lt / TPC;
        float2 *lds = lds_all + (sub + SUB0) * S::TILE_WORDS;
        const int tile = t0 + sub;
        const bool active = (S::NTB % NSUB == 0) || tile < S::NTB;
        const auto prepared = prep(tile, active);
        float2 v[16];
        if constexpr (PF) {
            // the first radix-16 butterfly consumes the prefetched values; the next tile's loads
            // then reuse their registers (issued before the rest of the transform)
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = nx[t];
            butterflies<NS, 16, 1, DIR>(v, 0, tw.n1);
            pin16(v);
            if (t0 + NSUB < t_end) fetch(t0 + NSUB, nx);
            __builtin_amdgcn_sched_barrier(0);
        } else {
            fetch(t0, v);
        }
        if constexpr (PF)
            stockham_rest<NS, NC, false, DIR, 16, 1>(v, lds, row, jt, tw.n1);
        else
            subfft<NS, NC, false, DIR>(v, lds, row, jt, tw.n1);
        __syncthreads();
        {
            constexpr int R = FinalShape<NS>::R;
#pragma unroll
            for (int q = 0; q < 16 / R; ++q)
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    // final_pos = A + (q TPC R + t Ns_last) with A per thread (no carries, as in
                    // stockham_rest); times NC (a multiple of 16) plus row keeps pad linear in it
                    constexpr int NsL = LastPass<NS>::Ns_last;
                    constexpr bool split = NC % 16 == 0 && ((NsL == 1 && R == 16) || (NsL % 16 == 0 && TPC % NsL == 0));
                    if constexpr (split) {
                        const int kk = jt & (NsL - 1);
                        const int A = (jt - kk) * R + kk;
                        const int C = q * TPC * R + t * NsL;
                        lds[A * (NC + NC / 16) + lds_pad(row) + C * (NC + NC / 16)] = v[q * R + t];
                    } else {
                        lds[lds_pad(final_pos<NS>(jt, q, t) * NC + row)] = v[q * R + t];
                    }
                }
        }
        __syncthreads();
        emit(lds, tile, active, prepared);
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

// zero the LDS fold histogram and the FP16-scale word (made visible by the first barrier of a unit)
__device__ __forceinline__ void fold_lds_init(double *hsum, unsigned int *hcnt, float *pw_lds, int nbins)
{
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        hsum[i] = 0.0;
        hcnt[i] = 0u;
    }
    if (threadIdx.x == 0) *pw_lds = 0.f;
}

// Epilogue variants of the dedispersion unit, fixed at compile time so that the plain path carries
// no per-sample tests: VAR_F16 adds the FP16-pair output and its block scale, VAR_TANGENT folds
// 2 Re(conj(ref) y) against a primal series instead of |y|^2.
constexpr int VAR_PLAIN = 0, VAR_F16 = 1, VAR_TANGENT = 2;

// Moves the LDS fold histogram of DM d into the global fold (threads t0, t0 + nthr, ... of the
// workgroup) and clears it for the next unit.
__device__ __forceinline__ void fold_flush(const InvArgs &a, int d, double *hsum, unsigned int *hcnt, int t0, int nthr)
{
    double *gs = a.fold_sum + (size_t)d * a.nbins;
    unsigned long long *gc = a.fold_cnt + (size_t)d * a.nbins;
    for (int i = t0; i < a.nbins; i += nthr) {
        const unsigned int c = hcnt[i];
        if (c) {
            unsafeAtomicAdd(&gs[i], hsum[i]);
            atomicAdd(&gc[i], (unsigned long long)c);
            hsum[i] = 0.0;
            hcnt[i] = 0u;
        }
    }
}

// Pass B of the dedispersion unit (DM d, block b): IDFT rows from the scratch slot (written in
// `layout`), then the kept half [M, L) to the outputs and into the fold histogram.
template <int N1, int N2, int VAR, bool PF, int NSUB, int SUB0 = 0>
__device__ __forceinline__ void inv_unit_pass_b(const InvArgs &a, int d, int b, float2 *lds_all, const Tw &tw,
                                                const float2 *__restrict__ scr, double *hsum, unsigned int *hcnt,
                                                float inv_scale, int layout = 0)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int L = S::L, M = L / 2;
    const bool fold = a.fold_sum != nullptr;
    constexpr bool f16 = (VAR & VAR_F16) != 0;
    constexpr bool tangent = (VAR & VAR_TANGENT) != 0;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    float2 *oc = a.out_c64 ? a.out_c64 + ((size_t)d * a.B + b) * M : nullptr;
    uint32_t *ou = a.out_u16 ? a.out_u16 + ((size_t)d * a.B + b) * M : nullptr;
    uint32_t *oh = f16 && a.out_f16 ? a.out_f16 + ((size_t)d * a.B + b) * M : nullptr;
    const uint16_t *bins = a.bins ? a.bins + (size_t)b * M : nullptr;
    const float2 *tref = tangent ? a.tangent_ref + ((size_t)d * a.B + b) * M : nullptr;
    // upper half of a tile's outputs (n1 >= N1/2): 8 consecutive f per thread, i.e. runs of
    // consecutive samples; the fold accumulates a run in registers and flushes on a bin change.
    // Runs of >= 8 consecutive samples: their 8 bins are one aligned 16-byte load, issued by prep.
    auto prep = [&](int tile, bool active) -> uint4 {
        constexpr int NC = S::NCB;
        uint4 b4 = make_uint4(0u, 0u, 0u, 0u);
        if (fold && NC >= 8 && active) {
            const int fbase = (N1 / 2) * NC + (tid & 255) * 8;
            const int n1 = fbase / NC, row = fbase % NC;
            b4 = *reinterpret_cast<const uint4 *>(bins + (tile * NC + row + N2 * n1 - M));
        }
        return b4;
    };
    auto emit = [&](const float2 *lds, int tile, bool active, const uint4 b4) {
        if (!active) return;
        constexpr int NC = S::NCB;
        const int lt = tid & 255;
        const int fbase = (N1 / 2) * NC + lt * 8;
        int cur = -1;
        double csum = 0.0;
        unsigned int ccnt = 0;
        uint16_t bv[8];
        if (fold && NC >= 8) {
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
            // fbase is a multiple of 8 and i < 8: no carry out of the low 4 bits, pad(f) = pad(fbase) + i
            float2 y = lds[lds_pad(fbase) + i];
            if (oc) st32(oc, (uint32_t)m, y);
            if (ou) st32(ou, (uint32_t)m, (enc_u16(y.y) << 16) | enc_u16(y.x));
            if constexpr (f16)
                if (oh) st32(oh, (uint32_t)m, enc_f16_pair(y, inv_scale));
            if (fold) {
                // primal: I = |y|^2. Tangent (y = d series / d DM, ref = series): the sample's
                // contribution is dI/dDM = 2 Re(conj(ref) y), kept or skipped as the primal was
                float I = y.x * y.x + y.y * y.y, v = I;
                if constexpr (tangent) {
                    const float2 r = ld32(tref, (uint32_t)m);
                    I = r.x * r.x + r.y * r.y;
                    v = 2.f * (r.x * y.x + r.y * y.y);
                }
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
                    csum += (double)v;
                    ccnt += 1;
                }
            }
        }
        if (fold && ccnt) {
            atomicAdd(&hsum[cur], csum);
            atomicAdd(&hcnt[cur], ccnt);
        }
    };
    pass_b<N1, N2, +1, PF, NSUB, SUB0>(scr, prep, emit, lds_all, tw, layout);
}

// One dedispersion unit (DM d, block b): gather * H -> IDFT_L -> keep [M, L) * 1/L -> outputs and
// fold. hsum/hcnt are the workgroup's LDS fold histogram (zero on entry and on exit), pw_lds the
// FP16-scale reduction word (zero on entry and on exit).
// With PF pass B prefetches its next scratch tile (32 VGPRs); with PFA the gather runs with the
// next tile's delays in flight (16 VGPRs).
template <int N1, int N2, bool RING32, int VAR, bool PF = false, int NSUB = SUBT, bool PFA = PF>
__device__ __forceinline__ void inv_unit(const InvArgs &a, int d, int b, float2 *lds_all, const Tw &tw,
                                         float2 *__restrict__ scr, double *hsum, unsigned int *hcnt, float *pw_lds)
{
    using S = Shape<N1, N2, NSUB>;
    constexpr int L = S::L;
    const bool fold = a.fold_sum != nullptr;
    constexpr bool f16 = (VAR & VAR_F16) != 0;
    // a.H (and dH/dDM) carry the 1/L of the inverse transform: a power of two, so the series is the
    // same bits as scaling after the transform, one multiply per sample cheaper
    // opaque thread index: keeps thread-dependent values from being hoisted out of the caller's
    // unit loop (where they would stay live, and spill, across units)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int gm = (int)((a.g0 + b) % a.FBS); // ring slot of this block
    const int FBS = a.FBS;
    const float2 *__restrict__ Hd = a.H + (size_t)d * L;
    const uint16_t *__restrict__ dl = a.delay + (size_t)d * L;
    const float2 *__restrict__ ring = a.ring;
    float power = 0.f; // sum |Y|^2 of this thread's bins (FP16 output scale, Parseval; Y includes 1/L)
    auto load = [&](int k) -> float2 {
        int src = gm - (int)ld32(dl, (uint32_t)k); // delays are < FBS, so one wrap at most
        if (src < 0) src += FBS;
        const float2 x = RING32 ? ld32(ring, (uint32_t)src * (uint32_t)L + (uint32_t)k) : ring[(size_t)src * L + k];
        const float2 y = cmul(x, ld32(Hd, (uint32_t)k));
        if constexpr (f16) power += y.x * y.x + y.y * y.y;
        return y;
    };
    if constexpr (PFA)
        pass_a_gather_dl<N1, N2, RING32, f16, NSUB>(ring, FBS, gm, Hd, dl, scr, lds_all, tw, power);
    else
        pass_a<N1, N2, +1, PF, NSUB>(load, scr, lds_all, tw);
    float inv_scale = 1.f, out_scale = 1.f;
    if constexpr (f16) {
        // block rms of the series over the L outputs: sqrt(sum |Y|^2) with Y = X H / L (Parseval);
        // scale = 2^floor(log2(rms))
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) power += __shfl_xor(power, off, 64);
        if ((tid & 63) == 0) atomicAdd(pw_lds, power);
        __syncthreads();
        const float tot = *pw_lds;
        const float rms = sqrtf(tot);
        int e2 = 0;
        (void)frexpf(rms, &e2); // rms = m * 2^e2, m in [0.5, 1)
        out_scale = rms > 0.f ? ldexpf(1.f, e2 - 1) : 1.f;
        inv_scale = 1.f / out_scale; // exact: a power of two
        if (tid == 0 && a.out_scale) a.out_scale[(size_t)d * a.B + b] = out_scale;
    }
    __syncthreads();
    inv_unit_pass_b<N1, N2, VAR, PF, NSUB>(a, d, b, lds_all, tw, scr, hsum, hcnt, inv_scale);
    __syncthreads();
    if (f16 && tid == 0) *pw_lds = 0.f; // next unit's reduction (ordered by the barriers)
    if (fold) {
        fold_flush(a, d, hsum, hcnt, tid, 256 * NSUB);
        __syncthreads();
    }
}


// Unit order of the persistent dedispersion kernels (a unit is one (DM, block) pair). Workgroup w
// visits the sequence positions first, first + step, ... below end; a position names a unit or, at
// the edges of the tile orders, nothing.
//   tiles: the G/8 workgroups of an XCD run dg DMs x (G/8)/dg consecutive blocks at a time, so each
//          DM's chirp and delays and each block's ring slots are shared in that XCD's L2;
//   XCD DM-/block-major: workgroups w, w+8, ... share an XCD and get consecutive virtual ids;
//   linear: workgroup w runs units w, w+G, ...
// Positions are 32-bit (plans keep num_dms x batch x 4 below 2^31) and wave-uniform, kept in scalar
// registers: wide or vector loop state is what the register allocator spills, and a spill reload at the
// loop head waits for every outstanding store of the wave.
struct UnitOrder {
    int dg, bg, j, tb;
    bool tiles, block_major;
    int first, end, step;
    __device__ UnitOrder(const InvArgs &a, int G, int w)
    {
        const int order = a.order & MSOSM_ORDER_MASK;
        dg = order == MSOSM_ORDER_XCD_TILE2    ? 2
             : order == MSOSM_ORDER_XCD_TILE4  ? 4
             : order == MSOSM_ORDER_XCD_TILE8  ? 8
             : order == MSOSM_ORDER_XCD_TILE16 ? 16
                                               : 0;
        tiles = dg && G % 8 == 0 && (G / 8) % dg == 0;
        const bool xcd = order != MSOSM_ORDER_LINEAR && G % 8 == 0;
        block_major = order == MSOSM_ORDER_XCD_BLOCK_MAJOR;
        bg = tiles ? (G / 8) / dg : 1;
        j = w / 8;
        tb = (a.B + bg - 1) / bg;
        const int td = tiles ? (a.D + dg - 1) / dg : 0;
        first = __builtin_amdgcn_readfirstlane(tiles ? w % 8 : xcd ? (w % 8) * (G / 8) + w / 8 : w);
        end = __builtin_amdgcn_readfirstlane(tiles ? td * tb : a.D * a.B);
        step = tiles ? 8 : G;
    }
    // the unit at position s; false when the position names none
    __device__ bool at(int s, const InvArgs &a, int &d, int &b) const
    {
        if (tiles) {
            d = (s / tb) * dg + j / bg;
            b = (s % tb) * bg + j % bg;
        } else if (block_major) {
            d = s % a.D;
            b = s / a.D;
        } else {
            d = s / a.B;
            b = s % a.B;
        }
        d = __builtin_amdgcn_readfirstlane(d);
        b = __builtin_amdgcn_readfirstlane(b);
        return d < a.D && b < a.B;
    }
    // the next unit from position s on (s moves past it); false at the end of the sequence
    __device__ bool next(int &s, const InvArgs &a, int &d, int &b) const
    {
        for (; s < end; s += step)
            if (at(s, a, d, b)) {
                s += step;
                return true;
            }
        return false;
    }
};

// Role-split dedispersion of square lengths (N1 = N2 = N, plain epilogue). Waves 0-7 (256-thread
// groups 0, 1) run pass A of the workgroup's unit j while waves 8-15 (groups 2, 3) run pass B of unit
// j-1 on the same scratch slot: units alternate the slot layout, so pass A writes each tile pair
// exactly where pass B has just read its own (slot_pos), and the memory-heavy gather and the
// slot read-back of one role overlap the other role's transforms. Both roles pass the same
// sequence of barriers: per phase (one unit of each role), STEPS tile steps of BPS barriers (pass A's
// sub-FFT exchanges and 2 more; pass B's exchanges and its sample-order write) and 2 at the phase end,
// between which the pass-B role moves its fold histogram out. A role without a unit in a phase (the
// pass-B role in the first, the pass-A role in the last) only passes the barriers.
template <int N, bool RING32>
__device__ __forceinline__ void inv_rs(const InvArgs &a, float2 *lds_all, const Tw &tw, float2 *__restrict__ scr,
                                       double *hsum, unsigned int *hcnt)
{
    using S = Shape<N, N, 2>;
    static_assert(S::NTA == S::NTB && S::NCA == S::NCB && S::NTA >= 4, "role split: square lengths, 4+ tiles");
    constexpr int STEPS = S::NTA / 2, L = S::L;
    // barriers of one tile step in either role: the sub-FFT's exchanges (two each) and 2 more
    constexpr int BPS = 2 * (PassCount<N>::value - 1) + 2;
    const UnitOrder o(a, gridDim.x, blockIdx.x);
    const int role = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 9);
    int s = o.first;
    int d = 0, b = 0;
    // the unit loops are unrolled by two, one body per slot layout, so that every slot address is
    // a compile-time function of the thread's indices (a runtime layout costs 8 % more VALU). The
    // pass-A role looks one unit ahead and starts its loads during the current unit's last tile step
    // (the ring and delay tables are read-only); the pass-B role cannot, since the first tile of its
    // next unit depends on every column that pass A stores in the same phase.
    bool have = o.next(s, a, d, b), first = true;
    if (role == 0) {
        int dd[16];
        float2 x[16];
        auto unit = [&](auto layout) -> bool {
            if (!have) {
                for (int i = 0; i < BPS * STEPS + 2; ++i) __syncthreads(); // the pass-B role's last unit
                return false;
            }
            int dn = 0, bn = 0;
            const bool next = o.next(s, a, dn, bn);
            const int gm = (int)((a.g0 + b) % a.FBS), gmn = (int)((a.g0 + bn) % a.FBS);
            pass_a_gather_pf<N, N, RING32, 2, 2>(a.ring, a.FBS, gm, a.H + (size_t)d * L, a.delay + (size_t)d * L,
                                                 scr, lds_all, tw, decltype(layout)::value, x, dd, first, next, gmn,
                                                 a.delay + (size_t)dn * L);
            __syncthreads(); // phase end (the pass-B role flushes its fold in between)
            __syncthreads();
            first = false;
            d = dn;
            b = bn;
            have = next;
            return true;
        };
        while (unit(IntC<0>{}) && unit(IntC<1>{})) {
        }
    } else {
        for (int i = 0; i < BPS * STEPS; ++i) __syncthreads(); // phase 0: pass A of the first unit
        __syncthreads();
        __syncthreads();
        const bool fold = a.fold_sum != nullptr;
        auto unit = [&](auto layout) -> bool {
            if (!have) return false;
            inv_unit_pass_b<N, N, VAR_PLAIN, true, 2, 2>(a, d, b, lds_all, tw, scr, hsum, hcnt, 1.f,
                                                         decltype(layout)::value);
            __syncthreads();
            if (fold) fold_flush(a, d, hsum, hcnt, (int)threadIdx.x - 512, 512);
            __syncthreads();
            have = o.next(s, a, d, b);
            return true;
        };
        while (unit(IntC<0>{}) && unit(IntC<1>{})) {
        }
    }
}

} // namespace msosm
