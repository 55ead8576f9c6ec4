// Register/LDS Stockham FFT building blocks for gfx950 (CDNA4, wave64).
//
// A sub-FFT of NS points (32..4096) is spread over NS/16 threads holding 16 points each; NCOL such
// sub-FFTs ("columns") form a 4096-point tile worked by 256 threads. Passes are radix-16 (one
// butterfly per thread) followed by one radix-2/4/8/16 pass; between passes the points go through
// LDS. Twiddles inside a pass come from one table read per butterfly (FP64-generated, rounded
// once) and its powers (at most 4 products deep).
//
// Sign convention: DIR = -1 computes sum_k x[k] e^{-2 pi i nk/N} (forward, as cuFFT/FFTW FORWARD),
// DIR = +1 computes the e^{+2 pi i nk/N} sum (inverse, unnormalised).
#pragma once

#include <hip/hip_runtime.h>

namespace msosm {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// a * (DIR * i)
template <int DIR>
__device__ __forceinline__ float2 mul_di(float2 a)
{
    return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// e^{DIR * 2 pi i m / 16}, m known at compile time
template <int DIR, int M>
__device__ __forceinline__ float2 w16()
{
    constexpr float C1 = 0.923879532511286756f, S1 = 0.382683432365089772f, C2 = 0.707106781186547524f;
    constexpr int m = M & 15;
    constexpr float c = m == 0 ? 1.f : m == 1 ? C1 : m == 2 ? C2 : m == 3 ? S1 : m == 4 ? 0.f : m == 5 ? -S1
                      : m == 6 ? -C2 : m == 7 ? -C1 : m == 8 ? -1.f : m == 9 ? -C1 : m == 10 ? -C2 : m == 11 ? -S1
                      : m == 12 ? 0.f : m == 13 ? S1 : m == 14 ? C2 : C1;
    constexpr float s = m == 0 ? 0.f : m == 1 ? S1 : m == 2 ? C2 : m == 3 ? C1 : m == 4 ? 1.f : m == 5 ? C1
                      : m == 6 ? C2 : m == 7 ? S1 : m == 8 ? 0.f : m == 9 ? -S1 : m == 10 ? -C2 : m == 11 ? -C1
                      : m == 12 ? -1.f : m == 13 ? -C1 : m == 14 ? -C2 : -S1;
    return make_float2(c, DIR * s);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2 &a, float2 &b)
{
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2 &v0, float2 &v1, float2 &v2, float2 &v3)
{
    float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = mul_di<DIR>(csub(v1, v3));
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, a3);
    v3 = csub(a1, a3);
}

// in-place DFT of v[0..R) with natural-order input and output
template <int DIR, int R>
__device__ __forceinline__ void dft(float2 *v)
{
    if constexpr (R == 2) {
        dft2<DIR>(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4<DIR>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // t = t0 + 2 t1, r = r0 + 4 r1
        dft4<DIR>(v[0], v[2], v[4], v[6]);
        dft4<DIR>(v[1], v[3], v[5], v[7]);
        v[3] = cmul(v[3], w16<DIR, 2>());
        v[5] = mul_di<DIR>(v[5]); // w16^4 = DIR * i: a swap and a sign, no multiplies
        v[7] = cmul(v[7], w16<DIR, 6>());
        dft2<DIR>(v[0], v[1]);
        dft2<DIR>(v[2], v[3]);
        dft2<DIR>(v[4], v[5]);
        dft2<DIR>(v[6], v[7]);
        // v[2 r0 + r1] holds V[r0 + 4 r1]
        float2 o[8];
#pragma unroll
        for (int r0 = 0; r0 < 4; ++r0)
#pragma unroll
            for (int r1 = 0; r1 < 2; ++r1) o[r0 + 4 * r1] = v[2 * r0 + r1];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = o[i];
    } else if constexpr (R == 16) {
        // t = t0 + 4 t1, r = r0 + 4 r1
#pragma unroll
        for (int t0 = 0; t0 < 4; ++t0) dft4<DIR>(v[t0], v[t0 + 4], v[t0 + 8], v[t0 + 12]);
        v[5] = cmul(v[5], w16<DIR, 1>());
        v[9] = cmul(v[9], w16<DIR, 2>());
        v[13] = cmul(v[13], w16<DIR, 3>());
        v[6] = cmul(v[6], w16<DIR, 2>());
        v[10] = mul_di<DIR>(v[10]);
        v[14] = cmul(v[14], w16<DIR, 6>());
        v[7] = cmul(v[7], w16<DIR, 3>());
        v[11] = cmul(v[11], w16<DIR, 6>());
        v[15] = cmul(v[15], w16<DIR, 9>());
#pragma unroll
        for (int r0 = 0; r0 < 4; ++r0) dft4<DIR>(v[4 * r0], v[4 * r0 + 1], v[4 * r0 + 2], v[4 * r0 + 3]);
        // v[4 r0 + r1] holds V[r0 + 4 r1]
        float2 o[16];
#pragma unroll
        for (int r0 = 0; r0 < 4; ++r0)
#pragma unroll
            for (int r1 = 0; r1 < 4; ++r1) o[r0 + 4 * r1] = v[4 * r0 + r1];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = o[i];
    }
}

// w[t] = w1^t for t in [0, R) with multiply chains of depth <= 4
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 w1, float2 *w)
{
    w[0] = make_float2(1.f, 0.f);
    w[1] = w1;
    if constexpr (R > 2) {
        w[2] = cmul(w1, w1);
        w[3] = cmul(w[2], w1);
    }
    if constexpr (R > 4) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[4], w1);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R > 8) {
        w[8] = cmul(w[4], w[4]);
#pragma unroll
        for (int t = 9; t < 16; ++t) w[t] = cmul(w[8], w[t - 8]);
    }
}

// One pad word after every 16: the radix-16 exchange writes with a stride of 16 elements and reads
// with a stride of 1, and the pad turns the strided side into a stride of 17 (no LDS bank conflicts
// for 64 banks x 4 B and 8-byte accesses; see the layout notes in kernels.hip).
__device__ __forceinline__ int lds_pad(int e) { return e + (e >> 4); }

// Index of element e of column col inside a tile's LDS region.
// COLFAST: lanes run along columns (pad(e) * NCOL + col); otherwise each column is a padded run.
template <int NS, int NCOL, bool COLFAST>
__device__ __forceinline__ int lds_index(int e, int col)
{
    if constexpr (COLFAST)
        return lds_pad(e) * NCOL + col;
    else
        return col * (NS + NS / 16) + lds_pad(e);
}

// lds_index of element A + C where pad(A + C) = pad(A) + C + C/16, i.e. where A + C has no carry out
// of the low 4 bits of A (C a multiple of 16, or A a multiple of 16). With C known at compile time
// the per-thread part pad(A) is computed once and each access is an immediate offset from it.
template <int NS, int NCOL, bool COLFAST>
__device__ __forceinline__ int lds_index_split(int padA, int C, int col)
{
    const int e = padA + C + (C >> 4);
    if constexpr (COLFAST)
        return e * NCOL + col;
    else
        return col * (NS + NS / 16) + e;
}

template <int NS, int NCOL, bool COLFAST>
struct TileLDS {
    static constexpr int ELEMS = (NS + NS / 16) * NCOL;
};

// Radix of the pass that starts at span Ns.
template <int NS, int Ns>
struct PassRadix {
    static constexpr int R = (NS / Ns) >= 16 ? 16 : (NS / Ns);
};

// Number of radix passes of an NS-point sub-FFT. Every pass after the first starts with an LDS
// exchange of two workgroup barriers.
template <int NS, int Ns = 1>
struct PassCount {
    static constexpr int value = 1 + PassCount<NS, Ns * PassRadix<NS, Ns>::R>::value;
};
template <int NS>
struct PassCount<NS, NS> {
    static constexpr int value = 0;
};
static_assert(PassCount<16>::value == 1 && PassCount<128>::value == 2 && PassCount<256>::value == 2 &&
                  PassCount<512>::value == 3 && PassCount<2048>::value == 3,
              "radix-16 passes, the last one of radix NS / 16^k");

// One pass: twiddle by e^{DIR 2 pi i t k/(Ns R)} (k = j mod Ns) and DFT_R, on the 16/R butterflies
// a thread owns (j = jt + q * NS/16). Twiddle table tw[m] = e^{+2 pi i m / NS}.
template <int NS, int R, int Ns, int DIR>
__device__ __forceinline__ void butterflies(float2 (&v)[16], int jt, const float2 *__restrict__ tw)
{
    constexpr int TPC = NS / 16;
    constexpr int NB = 16 / R;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        if constexpr (Ns > 1) {
            const int j = jt + q * TPC;
            const int k = j & (Ns - 1);
            float2 w1 = tw[k * (NS / (Ns * R))];
            if (DIR < 0) w1 = cconj(w1);
            // each power w^t once (R-2 products, depth <= 4: w^t = w^{t-hb} * w^{hb}, hb = highest bit)
            float2 w[R];
            w[1] = w1;
#pragma unroll
            for (int t = 2; t < R; ++t) {
                const int hb = (t & 8) ? 8 : (t & 4) ? 4 : (t & 2) ? 2 : 1;
                w[t] = (t == hb) ? cmul(w[hb / 2], w[hb / 2]) : cmul(w[hb], w[t - hb]);
            }
#pragma unroll
            for (int t = 1; t < R; ++t) v[q * R + t] = cmul(v[q * R + t], w[t]);
        }
        dft<DIR, R>(&v[q * R]);
    }
}

// Workgroup barrier that orders LDS accesses only: the fences cover the LDS address space, so the
// wave waits for its LDS operations but not for its global stores (which a full workgroup fence would
// make it wait for). For exchanges through LDS whose threads share nothing through global memory.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Barrier of an LDS exchange. WAVE: every point a wave reads was written by the same wave (its
// columns' regions are its own), and LDS instructions of one wave execute in order, so a
// wavefront-scope fence that keeps the compiler from moving accesses across it suffices. LDSB: a
// workgroup barrier over LDS only (lds_barrier).
template <bool WAVE, bool LDSB = false>
__device__ __forceinline__ void exchange_barrier()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else if constexpr (LDSB) {
        lds_barrier();
    } else {
        __syncthreads();
    }
}

// Stockham output position of element t of butterfly q after a pass (radix R, span Ns).
template <int NS, int R, int Ns>
__device__ __forceinline__ int out_pos(int jt, int q, int t)
{
    constexpr int TPC = NS / 16;
    const int j = jt + q * TPC;
    const int k = j & (Ns - 1);
    return (j - k) * R + k + t * Ns;
}

// Input position of element t of butterfly q for a pass of radix R.
template <int NS, int R>
__device__ __forceinline__ int in_pos(int jt, int q, int t)
{
    constexpr int TPC = NS / 16;
    return jt + q * TPC + t * (NS / R);
}

// Remaining passes after the one that ended at span Ns_done (radix R_done).
// On return, v holds the final outputs; element (q, t) is output index out_pos<NS, RL, NsL>.
template <int NS, int NCOL, bool COLFAST, int DIR, int R_done, int Ns_done, bool WAVE = false, bool LDSB = false>
__device__ __forceinline__ void stockham_rest(float2 (&v)[16], float2 *lds, int col, int jt,
                                              const float2 *__restrict__ tw)
{
    constexpr int Ns = Ns_done * R_done; // span at the start of the next pass
    if constexpr (Ns < NS) {
        constexpr int R = PassRadix<NS, Ns>::R;
        constexpr int TPC = NS / 16;
        exchange_barrier<WAVE, LDSB>(); // LDS free (previous readers done)
        // output position (j - k) R + k + t Ns, j = jt + q TPC, k = j mod Ns: split into a per-thread
        // part and a compile-time part without carries when k does not depend on q
        constexpr bool out_split = (Ns_done == 1 && R_done == 16) || (Ns_done % 16 == 0 && TPC % Ns_done == 0);
        if constexpr (out_split) {
            const int kk = jt & (Ns_done - 1);
            const int padA = lds_pad((jt - kk) * R_done + kk);
#pragma unroll
            for (int q = 0; q < 16 / R_done; ++q)
#pragma unroll
                for (int t = 0; t < R_done; ++t)
                    lds[lds_index_split<NS, NCOL, COLFAST>(padA, q * TPC * R_done + t * Ns_done, col)] =
                        v[q * R_done + t];
        } else {
#pragma unroll
            for (int q = 0; q < 16 / R_done; ++q)
#pragma unroll
                for (int t = 0; t < R_done; ++t)
                    lds[lds_index<NS, NCOL, COLFAST>(out_pos<NS, R_done, Ns_done>(jt, q, t), col)] = v[q * R_done + t];
        }
        exchange_barrier<WAVE, LDSB>();
        // input position jt + q TPC + t NS/R
        constexpr bool in_split = TPC % 16 == 0 && (NS / R) % 16 == 0;
        if constexpr (in_split) {
            const int padA = lds_pad(jt);
#pragma unroll
            for (int q = 0; q < 16 / R; ++q)
#pragma unroll
                for (int t = 0; t < R; ++t)
                    v[q * R + t] = lds[lds_index_split<NS, NCOL, COLFAST>(padA, q * TPC + t * (NS / R), col)];
        } else {
#pragma unroll
            for (int q = 0; q < 16 / R; ++q)
#pragma unroll
                for (int t = 0; t < R; ++t)
                    v[q * R + t] = lds[lds_index<NS, NCOL, COLFAST>(in_pos<NS, R>(jt, q, t), col)];
        }
        butterflies<NS, R, Ns, DIR>(v, jt, tw);
        stockham_rest<NS, NCOL, COLFAST, DIR, R, Ns, WAVE, LDSB>(v, lds, col, jt, tw);
    }
}

// Last-pass parameters for a sub-FFT of NS points.
template <int NS, int Ns = 1>
struct LastPass {
    static constexpr int R = PassRadix<NS, Ns>::R;
    static constexpr int Ns_last = (Ns * R == NS) ? Ns : LastPass<NS, Ns * R>::Ns_last;
    static constexpr int R_last = (Ns * R == NS) ? R : LastPass<NS, Ns * R>::R_last;
};
template <int NS>
struct LastPass<NS, NS> {
    static constexpr int R = 1, Ns_last = 1, R_last = 1;
};

// Full sub-FFT. On entry v[t] = x[jt + t * NS/16] (t in [0,16)). On exit element (q, t) of v is
// output index final_pos<NS>(jt, q, t). WAVE: the NS/16 threads of every column are lanes of one
// wave and the caller has made the columns' LDS regions free (exchange_barrier). LDSB: see
// exchange_barrier.
template <int NS, int NCOL, bool COLFAST, int DIR, bool WAVE = false, bool LDSB = false>
__device__ __forceinline__ void subfft(float2 (&v)[16], float2 *lds, int col, int jt, const float2 *__restrict__ tw)
{
    butterflies<NS, 16, 1, DIR>(v, jt, tw);
    stockham_rest<NS, NCOL, COLFAST, DIR, 16, 1, WAVE, LDSB>(v, lds, col, jt, tw);
}

template <int NS>
__device__ __forceinline__ int final_pos(int jt, int q, int t)
{
    return out_pos<NS, LastPass<NS>::R_last, LastPass<NS>::Ns_last>(jt, q, t);
}

template <int NS>
struct FinalShape {
    static constexpr int R = LastPass<NS>::R_last; // elements per butterfly in the last pass
};

} // namespace msosm
