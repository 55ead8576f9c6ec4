// Dedispersion units without a scratch slot, for FFT lengths L = R * S with S = 2^14 and R = 1, 2, 4.
//
// A unit (DM d, block b) computes y = IDFT_L(Y), Y[k] = X_{b - delay(k)}[k] H[k], and keeps y[M + m],
// m in [0, M), M = L / 2. Decimation in time by R splits the inverse transform into R transforms of
// S points over the bins k = R k' + r:
//   C_r[j] = sum_{k'} Y[R k' + r] e^{+2 pi i k' j / S},   y[n] = sum_r e^{+2 pi i r n / L} C_r[n mod S].
// Each C_r is a four-step 128 x 128 transform that lives in the workgroup's LDS (128 KiB), and the kept
// outputs are accumulated in registers across r (M / 1024 complex values per thread):
//   R = 1: y[M + m] = C_0[S/2 + m];
//   R = 2: y[M + j] = C_0[j] - W^j C_1[j];
//   R = 4: y[M + j] = sum_r (-1)^r W^{rj} C_r[j],  y[M + S + j] = sum_r (-i)^r W^{rj} C_r[j]
// with W = e^{+2 pi i / L} and j in [0, S). The four-step intermediate never leaves the CU: the launch
// reads the gathered spectra (ring, chirp, delays) and writes only its outputs.
//
// The launch reads its own copies of the ring, with bin k at position P(k) = (k mod R) S + k div R
// (perm_pos), so that the gather of C_r reads contiguous runs, of the chirp, in the order its threads
// use it (acc_chirp_pos: eight 1 KiB runs per wave and transform, loaded straight into LDS), and of the
// delays as bytes in thread order (one 16-byte load per thread and transform).
#pragma once
#include "msosm_device.h"

namespace msosm {

constexpr int ACC_S = 1 << 14, ACC_N = 128; // sub-transform length and its four-step split
constexpr int ACC_XS = ACC_N + 8;           // row stride of the LDS transpose (conflict-free both ways)
constexpr int ACC_AREA = 4 * TileWords<ACC_N>::V; // float2 words: tile exchanges / transpose / fold buffer
static_assert(ACC_AREA >= ACC_N * ACC_XS, "the transpose fits the tile area");

// Twiddle tables of an accumulate launch in LDS: the 128-point sub-transform table and the split
// W_L^m = hi[m >> s] lo[m & (2^s - 1)].
struct AccTw {
    const float2 *t128, *lo, *hi;
    int s;
};

template <int R>
struct AccShape {
    static constexpr int L = R * ACC_S, M = L / 2, MPT = M / WG; // kept outputs per thread
    static constexpr int FOLD_X = WG + 1;                        // fold buffer row stride (floats)
    static constexpr int TW_WORDS = 128 + 256 + 256;             // t128, lo (<= 256), hi (<= 256)
    static constexpr int SLOT_BYTES = 2 * 256 * 4;               // two ring-offset tables (this unit, next)
    static constexpr int LDS_BYTES = ACC_AREA * 8 + TW_WORDS * 8 + SLOT_BYTES;
    static_assert(MPT * FOLD_X <= ACC_AREA * 2, "fold buffer fits");
};

// the tables of the plan (one-length layout: n1, n2, lo, hi with s = (log2 L + 1) / 2) into LDS; with
// Q > 1 sub-units, the plan's tables of the sub-unit length
template <int R, int Q = 1>
__device__ __forceinline__ AccTw acc_load_twiddles(float2 *lds_all, const InvArgs &a)
{
    using A = AccShape<R>;
    float2 *t = lds_all + ACC_AREA;
    float2 *lo = t + 128, *hi = lo + 256;
    // a 128-point table: the plan's N2 table has 128 entries for L <= 2^15, its N1 table 256 at 2^16
    const float2 *src128 = Q > 1 ? a.sub_128 : R == 4 ? a.tw_n1 : a.tw_n2;
    constexpr int st = Q == 1 && R == 4 ? 2 : 1;
    const int s = Q > 1 ? a.sub_s : a.tw_s, nlo = 1 << s, nhi = A::L >> s;
    const float2 *src_lo = Q > 1 ? a.sub_lo : a.tw_lo, *src_hi = Q > 1 ? a.sub_hi : a.tw_hi;
    for (int i = threadIdx.x; i < 128; i += blockDim.x) t[i] = src128[i * st];
    for (int i = threadIdx.x; i < nlo; i += blockDim.x) lo[i] = src_lo[i];
    for (int i = threadIdx.x; i < nhi; i += blockDim.x) hi[i] = src_hi[i];
    __syncthreads();
    return AccTw{t, lo, hi, s};
}

// Ring offsets of block b's unit: tab[dd] = ((g0 + b - dd) mod FBS) L for delays dd < 256 (threads 0..255)
template <int R>
__device__ __forceinline__ void acc_slot_table(const InvArgs &a, int b, uint32_t *tab)
{
    if (threadIdx.x < 256) {
        int sl = (int)((a.g0 + b - (long long)threadIdx.x) % a.FBS);
        if (sl < 0) sl += a.FBS;
        tab[threadIdx.x] = (uint32_t)sl * (uint32_t)(R * ACC_S);
    }
}

__device__ __forceinline__ float2 acc_wl(const AccTw &tw, int m)
{
    return cmul(tw.hi[m >> tw.s], tw.lo[m & ((1 << tw.s) - 1)]);
}

// LDS rows of wave w in the pass-B exchange layout (8 rows of ACC_N + ACC_N / 16 words): where the
// wave's chirp values are staged
__device__ __forceinline__ float2 *acc_wave_rows(float2 *lds_all, int w)
{
    return lds_all + (w >> 2) * TileWords<ACC_N>::V + (w & 3) * 8 * (ACC_N + ACC_N / 16);
}

// Chirp values of transform r for wave w (acc_chirp_pos) from global memory straight into LDS: eight
// 16-byte loads per lane, lane l of load i landing at rows + 128 i + 2 l; complete once the wave's
// vector memory counter drains
__device__ __forceinline__ void acc_rows_from(const float2 *src, int step, float2 *rows)
{
#pragma unroll
    for (int i = 0; i < 8; ++i)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + step * i),
                                         (__attribute__((address_space(3))) void *)(rows + 128 * i), 16, 0, 0);
}

__device__ __forceinline__ void acc_chirp_to_lds(const float2 *Hr, float2 *rows, int w, int lane)
{
    acc_rows_from(Hr + 1024 * w + 2 * lane, 128, rows);
}

// The 16 byte delays of this thread for transform r of DM d: the bytes at r S + 16 p (p = k1 + 128 jtA),
// the t-th one for bin position r S + p + 1024 t
template <int R>
__device__ __forceinline__ uint4 acc_delays(const InvArgs &a, int d, int r)
{
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lt = tid & 255, p = (tid >> 8) * 32 + (lt & 31) + 128 * (lt >> 5);
    return *reinterpret_cast<const uint4 *>(a.delay_acc + (size_t)d * (R * ACC_S) + (size_t)r * ACC_S + 16u * (uint32_t)p);
}

// Transform index of acc_unit's body: V >= 0 known at compile time, V = -1 a run-time r
template <int V>
struct AccR {
    static constexpr int value = V;
};

// One dedispersion unit in the accumulate form. hsum/hcnt: the LDS fold histogram (zero on entry and
// exit); pw_lds: the FP16-scale word (zero on entry and exit).
// dl holds this unit's first delay words on entry and, on return, those of the next unit (DM dn, block
// bn, when next): each transform's delays are loaded one transform ahead. slot: this unit's ring offsets
// (acc_slot_table) on entry; the next unit's are written to slot_next before the unit's last barrier.
//
// Q > 1 sub-units (lengths L = Q 2^16, R = 4): the unit (d, b) is split by decimation in frequency into Q
// sub-units c, each the unit above of length L' = L / Q on the spectrum
//   X_c[k'] = sum_s Y[k' + s L'] e^{+2 pi i c (k' + s L') / L},   k' in [0, L'),
// whose kept outputs [L'/2, L') are y[M + Q u + c], u in [0, L'/2) (M = L / 2). The gathering pass
// (acc_pre_kernel) has written X_c in the chirp's order (a.xpre), so each transform loads its inputs
// into LDS as the chirp is loaded: no delays, ring offsets or chirps here, and dl and slot are unused. Sub-unit
// c writes every Q-th sample of the block; with sync (the unit's arrival counter, zero before its first
// sub-unit arrives) the workgroup waits, for at most ACC_SYNC_POLLS polls, until all Q sub-units of the
// unit have finished their transforms, so that the Q workgroups (on one XCD) write each cache line of
// the series at about the same time and the line leaves the L2 whole. The outputs do not depend on it.
// OUTK = ACC_OUT_C64 or ACC_OUT_U16 (sub-units with that one output and no fold, ONE_OUT): the item's first
// spectrum arrives in registers (xv, loaded by the previous item or the kernel), and the next item's
// (xn; after the last item this item's own again) is loaded into xv just before this item's output
// stores, so that the next
// item's first transform waits for those loads only and not for the stores (a wave's loads and stores
// share its memory counter, which the compiler's waits count precisely for register loads); the
// unit's barriers order LDS only, so the stores stay in flight through that transform.
constexpr int ACC_SYNC_POLLS = 96;
enum { ACC_OUT_ANY = 0, ACC_OUT_C64 = 1, ACC_OUT_U16 = 2 }; // outputs of a launch known at compile time

// The unit's workgroup barriers guard LDS only. LDSB: barriers that do not wait for the wave's global
// stores, so that one item's output stores stay in flight through the next item's first transform.
template <bool LDSB>
__device__ __forceinline__ void acc_barrier()
{
    if constexpr (LDSB)
        lds_barrier();
    else
        __syncthreads();
}

// The first spectrum of a sub-unit in registers, in the order of acc_unit's first transform: the row
// words of wave w, lane l (positions 1024 w + 128 i + 2 l and + 1)
__device__ __forceinline__ void acc_x0_load(const float2 *X, float2 (&xv)[16])
{
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const float4 *src = reinterpret_cast<const float4 *>(X + 1024 * (tid >> 6) + 2 * (tid & 63));
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float4 h = src[64 * i];
        xv[2 * i] = make_float2(h.x, h.y);
        xv[2 * i + 1] = make_float2(h.z, h.w);
    }
}

template <int R, bool RING32, int VAR, int Q = 1, int OUTK = ACC_OUT_ANY>
__device__ __forceinline__ void acc_unit(const InvArgs &a, int d, int b, int c, float2 *lds_all, const AccTw &tw,
                                         double *hsum, unsigned int *hcnt, float *pw_lds, uint4 &dl, bool next, int dn,
                                         int bn, const uint32_t *slot, uint32_t *slot_next,
                                         unsigned int *sync, const float2 *xn, float2 (&xv)[16])
{
    constexpr bool ONE_OUT = OUTK != ACC_OUT_ANY;
    static_assert(!ONE_OUT || (Q > 1 && R == 4), "one-output sub-units");
    using A = AccShape<R>;
    constexpr int L = A::L, M = A::M, S = ACC_S, MPT = A::MPT;
    constexpr bool f16 = (VAR & VAR_F16) != 0;
    constexpr bool tangent = (VAR & VAR_TANGENT) != 0;
    static_assert(Q == 1 || (R == 4 && !f16 && !tangent), "sub-units: R = 4, plain, uint16 or fold epilogues");
    constexpr int MB = Q * M; // samples per block
    const bool fold = !ONE_OUT && a.fold_sum != nullptr;
    const float2 *__restrict__ Hd = a.H_p + (size_t)d * L;
    const float2 *__restrict__ ring = a.ring_p;
    // sub-units: this sub-unit's spectrum (transform r at Xs + r S, in the chirp's order)
    const float2 *__restrict__ Xs = Q == 1 ? nullptr : a.xpre + (((size_t)d * a.B + b) * Q + c) * (size_t)L;
    float power = 0.f;
    // R = 4: transform 0 has its own code (its combine factors are 1: the accumulators start as its
    // outputs) and one loop runs r = 1..3; R = 1, 2: one loop from zeroed accumulators (a separate
    // transform 0 costs registers there, and a fully unrolled R = 4 spills heavily)
    constexpr bool peel0 = R == 4;
    float2 acc0[R == 1 ? 8 : 16], acc1[R == 4 ? 16 : 1];
    if constexpr (!peel0) {
#pragma unroll
        for (int e = 0; e < (R == 1 ? 8 : 16); ++e) acc0[e] = make_float2(0.f, 0.f);
    }
    // transform r: gather, C_r, and its share of the kept outputs (RC = r, or -1 for a run-time r)
    auto transform = [&](auto rc_, int r) {
        constexpr int RC = decltype(rc_)::value;
        // per-thread indices from an opaque thread index, so that their addresses are formed here and
        // not hoisted out of the loop (where they would stay live beside the accumulators)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int sub = tid >> 8, lt = tid & 255;
        const int col = lt & 31, jtA = lt >> 5; // pass A: column k1 = 32 sub + col, rows jtA + 8 t
        const int jtB = lt & 7, rowB = lt >> 3; // pass B: row n2 = 32 sub + rowB, columns jtB + 8 t
        const int k1 = sub * 32 + col, n2B = sub * 32 + rowB;
        float2 *tile = lds_all + sub * TileWords<ACC_N>::V;
        // gather of C_r's inputs: bins R (k1 + 128 (jtA + 8 t)) + r at positions r S + k1 + 128 jtA + 1024 t.
        // The ring values load into the transform's registers; the chirp is in the wave's LDS rows, put
        // there during the previous transform's combine or, for r = 0, requested here after the ring-offset
        // lookups (which it would otherwise hold up: an LDS read after an LDS-bound load waits for it), in
        // the same round trip as the ring. The next transform's delays follow.
        const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63; // wv: wave-uniform (SGPRs)
        float2 *rows = acc_wave_rows(lds_all, wv);
        float2 v[16];
        if constexpr (Q == 1) {
            const uint32_t p0 = (uint32_t)(r * S + k1 + 128 * jtA);
            const uint32_t dw[4] = {dl.x, dl.y, dl.z, dl.w};
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const uint32_t p = p0 + 1024u * t;
                const uint32_t off = slot[(dw[t >> 2] >> (8 * (t & 3))) & 255u] + p;
                v[t] = RING32 ? ld32(ring, off) : ring[(size_t)off];
            }
            if (RC == 0 || (RC < 0 && r == 0)) acc_chirp_to_lds(Hd, rows, wv, lane);
            if (r + 1 < R)
                dl = acc_delays<R>(a, d, r + 1);
            else if (next)
                dl = acc_delays<R>(a, dn, 0);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            const float4 *hrow = reinterpret_cast<const float4 *>(rows) + lane;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float4 h = hrow[64 * i];
                const float2 hv[2] = {make_float2(h.x, h.y), make_float2(h.z, h.w)};
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    float2 &y = v[2 * i + e];
                    y = cmul(y, hv[e]);
                    if constexpr (f16) power += y.x * y.x + y.y * y.y;
                }
            }
        } else {
            // the sub-unit spectrum from the gathering pass, in the chirp's order: in the wave's LDS rows,
            // put there during the previous transform's combine or, for r = 0, requested here (one
            // output: in registers already)
            if constexpr (ONE_OUT && RC == 0) {
#pragma unroll
                for (int t = 0; t < 16; ++t) v[t] = xv[t];
            } else {
                if (RC == 0 || (RC < 0 && r == 0)) acc_chirp_to_lds(Xs, rows, wv, lane);
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                const float4 *hrow = reinterpret_cast<const float4 *>(rows) + lane;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float4 h = hrow[64 * i];
                    v[2 * i] = make_float2(h.x, h.y);
                    v[2 * i + 1] = make_float2(h.z, h.w);
                }
            }
        }
        // pass A: columns over k2, then W_L^{n2 (R k1 + r)} (the four-step twiddle times W_L^{r n2})
        subfft<ACC_N, 32, true, +1, false, ONE_OUT>(v, tile, col, jtA, tw.t128);
        {
            const int kk = R * k1 + r;
            const float2 c = acc_wl(tw, 16 * kk); // n2 steps by 16 along t
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                float2 w = acc_wl(tw, (jtA + 8 * q) * kk);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    v[q * 8 + t] = cmul(v[q * 8 + t], w);
                    if (t + 1 < 8) w = cmul(w, c);
                }
            }
        }
        acc_barrier<ONE_OUT>(); // every tile's exchange reads are done: the area becomes the transpose
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int t = 0; t < 8; ++t) lds_all[(jtA + 8 * q + 16 * t) * ACC_XS + k1] = v[q * 8 + t];
        acc_barrier<ONE_OUT>();
#pragma unroll
        for (int t = 0; t < 16; ++t) v[t] = lds_all[n2B * ACC_XS + jtB + 8 * t];
        acc_barrier<ONE_OUT>(); // the transpose is read: the area becomes the tile exchanges again
        // pass B: rows over k1; element (q, t) is n1 = jtB + 8 q + 16 t, j = n2B + 128 n1. The 8 threads
        // of a row are consecutive lanes and a row's exchange region (row-major, 136 words) is its own,
        // so the exchanges need no workgroup barrier
        subfft<ACC_N, 32, false, +1, true>(v, tile, rowB, jtB, tw.t128);
        // accumulate T_r[j] = W_L^{128 r n1} C'_r[j] (the factor W_L^{r n2} rode in pass A's twiddle):
        //   R = 2: acc0 += (-1)^r T;  R = 4: acc0 += (-1)^r T, acc1 += (-i)^r T;  R = 1: keep n1 >= 64.
        // One branch-free path for every r (r is wave-uniform): the factors come from the LDS table.
        if constexpr (R == 1) {
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int t = 4; t < 8; ++t) acc0[q * 4 + t - 4] = v[q * 8 + t];
        } else {
            // W_L^{128 r n1} for n1 = jtB + 8 q + 16 t: a per-thread start and two wave-uniform steps
            // (chains of at most 8 products); all 1 at r = 0
            float2 cq = make_float2(1.f, 0.f), ct = cq, w = cq;
            if constexpr (RC != 0) {
                cq = acc_wl(tw, 1024 * r);
                ct = acc_wl(tw, 2048 * r);
                w = acc_wl(tw, 128 * r * jtB);
            }
            // the next transform's chirp (sub-units: its spectrum) into this wave's rows: their pass-B
            // exchange reads are complete,
            // no other wave touches them before the next pass A's first barrier, and the combine below reads
            // no LDS. Without a branch on r (one costs the register allocator dearly) the last transform
            // requests its own chirp again; that request drains before the rows become the fold buffer.
            // The scheduling barrier keeps the requests ahead of the combine.
            // Sub-units have nothing to request after the last transform (their spectra are read once): that
            // request reads one 16-byte word into every row word instead (the rows are not read again
            // before the next transform's own request), where a repeated spectrum would miss the L2s.
            {
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                const float2 *src = (Q == 1 ? Hd : Xs) + (size_t)(r + 1 < R ? r + 1 : r) * S;
                const bool real = Q == 1 || r + 1 < R;
                acc_rows_from(real ? src + 1024 * wv + 2 * lane : src, real ? 128 : 0, rows);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int e = q * 8 + t;
                    if constexpr (RC == 0) {
                        // T_0 = C'_0: the accumulators start here
                        acc0[e] = v[e];
                        if constexpr (R == 4) acc1[e] = v[e];
                    } else if constexpr (RC > 0) {
                        // (-1)^r and (-i)^r known: signs and a swap
                        const float2 y = cmul(v[e], q ? cmul(w, cq) : w);
                        if constexpr (RC & 1) {
                            acc0[e].x -= y.x;
                            acc0[e].y -= y.y;
                        } else {
                            acc0[e].x += y.x;
                            acc0[e].y += y.y;
                        }
                        if constexpr (R == 4) {
                            if constexpr (RC == 1) { // -i y
                                acc1[e].x += y.y;
                                acc1[e].y -= y.x;
                            } else if constexpr (RC == 2) {
                                acc1[e].x -= y.x;
                                acc1[e].y -= y.y;
                            } else { // i y
                                acc1[e].x -= y.y;
                                acc1[e].y += y.x;
                            }
                        }
                    } else {
                        // one path for every run-time r (a branch on r here costs the register allocator
                        // dearly): (-1)^r and (-i)^r = cr + i sr as factors
                        const float sg = (r & 1) ? -1.f : 1.f;
                        const float cr = (r & 1) ? 0.f : (r & 2) ? -1.f : 1.f;
                        const float sr = (r & 1) ? ((r & 2) ? 1.f : -1.f) : 0.f;
                        const float2 y = cmul(v[e], q ? cmul(w, cq) : w);
                        acc0[e].x += sg * y.x;
                        acc0[e].y += sg * y.y;
                        if constexpr (R == 4) {
                            acc1[e].x += cr * y.x - sr * y.y;
                            acc1[e].y += cr * y.y + sr * y.x;
                        }
                    }
                }
                if constexpr (RC != 0)
                    if (t + 1 < 8) w = cmul(w, ct);
            }
        }
        // no barrier here: the next transform's first cross-wave LDS write follows pass A's first
        // workgroup barrier, which every wave reaches only after this combine and its next chirp reads
    };
    if constexpr (peel0) transform(AccR<0>{}, 0);
#pragma unroll 1
    for (int r = peel0 ? 1 : 0; r < R; ++r) transform(AccR<-1>{}, r);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); // the last chirp request has landed
    if constexpr (Q > 1) {
        if (sync && threadIdx.x == 0) {
            __hip_atomic_fetch_add(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int i = 0; i < ACC_SYNC_POLLS &&
                            __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < (unsigned int)Q;
                 ++i)
                __builtin_amdgcn_s_sleep(8);
        }
    }
    __syncthreads(); // every wave's pass B is done: the area becomes the fold buffer
    if constexpr (ONE_OUT) {
        // the next item's first spectrum, ahead of the output stores (after the last item, this item's
        // again: an unconditional load leaves xv dead through the transforms)
        acc_x0_load(xn ? xn : Xs, xv);
    }

    // outputs: element e of acc0 is sample m = j (R = 2, 4) or j - S/2 (R = 1), acc1 is m = j + S
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int jtB = tid & 7, n2B = (tid >> 8) * 32 + ((tid & 255) >> 3);
    float inv_scale = 1.f;
    if constexpr (f16) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) power += __shfl_xor(power, off, 64);
        if ((tid & 63) == 0) atomicAdd(pw_lds, power);
        __syncthreads();
        const float rms = sqrtf(*pw_lds);
        int e2 = 0;
        (void)frexpf(rms, &e2);
        const float out_scale = rms > 0.f ? ldexpf(1.f, e2 - 1) : 1.f;
        inv_scale = 1.f / out_scale;
        if (tid == 0 && a.out_scale) a.out_scale[(size_t)d * a.B + b] = out_scale;
    }
    float2 *oc = OUTK == ACC_OUT_C64 || (OUTK == ACC_OUT_ANY && a.out_c64) ? a.out_c64 + ((size_t)d * a.B + b) * MB : nullptr;
    uint32_t *ou = OUTK == ACC_OUT_U16 || (OUTK == ACC_OUT_ANY && a.out_u16) ? a.out_u16 + ((size_t)d * a.B + b) * MB : nullptr;
    uint32_t *oh = f16 && a.out_f16 ? a.out_f16 + ((size_t)d * a.B + b) * MB : nullptr;
    const float2 *tref = tangent ? a.tangent_ref + ((size_t)d * a.B + b) * MB : nullptr;
    float *fb = reinterpret_cast<float *>(lds_all); // fold buffer: sample m at (m % MPT) FOLD_X + m / MPT
    // the fold bins of the samples this thread folds, [MPT tid, MPT tid + MPT), requested before the
    // outputs are written so that their latency overlaps that
    uint4 bw[MPT / 8];
    if (fold) {
        // sub-units: the bins of sub-unit c's samples are contiguous (launch_bins_sub)
        const uint4 *bins4 = reinterpret_cast<const uint4 *>(a.bins + ((size_t)b * Q + c) * M + (size_t)MPT * tid);
#pragma unroll
        for (int i = 0; i < MPT / 8; ++i) bw[i] = bins4[i];
    }
    // fold buffer position of sample m = n2B + 128 n1 + e (e = 0, S or -S/2, all multiples of MPT, as 128
    // is): (m mod MPT) FOLD_X + m div MPT = fb0 + (128 / MPT) n1 + e / MPT, a per-thread base and an
    // offset known at compile time
    static_assert(128 % MPT == 0 && (S / 2) % MPT == 0, "fold buffer positions split");
    const int fb0 = (n2B % MPT) * A::FOLD_X + n2B / MPT;
    auto out1 = [&](int m, int fbi, float2 y) {
        const uint32_t mo = Q == 1 ? (uint32_t)m : (uint32_t)(Q * m + c); // sample index in the block
        if (OUTK == ACC_OUT_C64 || (OUTK == ACC_OUT_ANY && oc)) st32(oc, mo, y);
        if (OUTK == ACC_OUT_U16 || (OUTK == ACC_OUT_ANY && ou)) st32(ou, mo, (enc_u16(y.y) << 16) | enc_u16(y.x));
        if constexpr (f16)
            if (oh) st32(oh, mo, enc_f16_pair(y, inv_scale));
        if (fold) {
            // primal: |y|^2; tangent: 2 Re(conj(ref) y), NaN where the primal sample is skipped (|ref| = 0)
            float v = y.x * y.x + y.y * y.y;
            if constexpr (tangent) {
                const float2 rf = ld32(tref, mo);
                const float I = rf.x * rf.x + rf.y * rf.y;
                v = I != 0.0f ? 2.f * (rf.x * y.x + rf.y * y.y) : __builtin_nanf("");
            }
            fb[fbi] = v;
        }
    };
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int n1 = jtB + 8 * q + 16 * t, j = n2B + 128 * n1;
            const int fbi = fb0 + (128 / MPT) * jtB + (128 / MPT) * (8 * q + 16 * t);
            if constexpr (R == 1) {
                if (t >= 4) out1(j - S / 2, fbi - (S / 2) / MPT, acc0[q * 4 + t - 4]);
            } else {
                out1(j, fbi, acc0[q * 8 + t]);
                if constexpr (R == 4) out1(j + S, fbi + S / MPT, acc1[q * 8 + t]);
            }
        }
    if (fold) {
        __syncthreads();
        // thread tid folds samples [MPT tid, MPT tid + MPT) with runs of equal bin merged in registers
        uint16_t bv[MPT];
#pragma unroll
        for (int i = 0; i < MPT / 8; ++i) {
            const uint32_t w[4] = {bw[i].x, bw[i].y, bw[i].z, bw[i].w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                bv[8 * i + 2 * k] = (uint16_t)(w[k] & 0xFFFFu);
                bv[8 * i + 2 * k + 1] = (uint16_t)(w[k] >> 16);
            }
        }
        int cur = -1;
        double csum = 0.0;
        unsigned int ccnt = 0;
#pragma unroll
        for (int i = 0; i < MPT; ++i) {
            const float v = fb[i * A::FOLD_X + tid];
            const bool keep = tangent ? !__builtin_isnan(v) : v != 0.0f;
            if (keep) {
                const int bin = (int)bv[i];
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
        if (ccnt) {
            atomicAdd(&hsum[cur], csum);
            atomicAdd(&hcnt[cur], ccnt);
        }
    }
    acc_barrier<ONE_OUT>();
    if (f16 && tid == 0) *pw_lds = 0.f;
    if constexpr (Q == 1)
        if (next) acc_slot_table<R>(a, bn, slot_next);
    if (fold) fold_flush(a, d, hsum, hcnt, tid, WG);
    acc_barrier<ONE_OUT>();
}

} // namespace msosm
