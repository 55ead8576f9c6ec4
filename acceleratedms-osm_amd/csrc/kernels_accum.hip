// Persistent dedispersion kernels of the accumulate form: one workgroup per CU walks (DM, block) units
// in the plan's unit order (FFT lengths 2^14, 2^15, 2^16), or (DM, block, sub-unit) items after a
// gathering pass (lengths 2^17, 2^18: Q = L / 2^16 sub-units of 2^16 points); each unit's inverse
// transform stays in the CU (LDS and registers), so the launch reads the gathered spectra and writes
// only its outputs.
#include "msosm_accum.h"

namespace msosm {

template <int R, bool RING32, int VAR>
__global__ void __launch_bounds__(WG) acc_kernel(InvArgs a)
{
    using A = AccShape<R>;
    extern __shared__ float2 lds_all[];
    uint32_t *slots = reinterpret_cast<uint32_t *>(lds_all + ACC_AREA + A::TW_WORDS); // two tables of 256
    const UnitOrder o(a, gridDim.x, blockIdx.x);
    int s = o.first;
    int d = 0, b = 0;
    bool have = o.next(s, a, d, b); // uniform over the workgroup
    if (have) acc_slot_table<R>(a, b, slots); // made visible by the twiddle load's barrier
    const AccTw tw = acc_load_twiddles<R>(lds_all, a);
    double *hsum = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_all) + A::LDS_BYTES);
    unsigned int *hcnt = reinterpret_cast<unsigned int *>(hsum + a.nbins);
    float *pw_lds = reinterpret_cast<float *>(hcnt + a.nbins + (a.nbins & 1));
    fold_lds_init(hsum, hcnt, pw_lds, a.fold_sum ? a.nbins : 0);
    // the workgroup's units in the plan's order, each unit's first delays and ring offsets prepared
    // during the previous one
    uint4 dl = have ? acc_delays<R>(a, d, 0) : make_uint4(0u, 0u, 0u, 0u);
    int cur = 0;
    float2 xv[16]; // unused by one-unit launches
#pragma unroll 1
    while (have) {
        int dn = 0, bn = 0;
        const bool next = o.next(s, a, dn, bn);
        acc_unit<R, RING32, VAR>(a, d, b, 0, lds_all, tw, hsum, hcnt, pw_lds, dl, next, dn, bn, slots + 256 * cur,
                                 slots + 256 * (cur ^ 1), nullptr, nullptr, xv);
        cur ^= 1;
        d = dn;
        b = bn;
        have = next;
    }
}

// Gathering pass of the sub-unit accumulate launch (L = Q 2^16): one workgroup per (DM d, chunk i of 8192
// sub-spectrum bins), over every block b of the push. For each bin k' it reads the Q bins
// k = k' + s 2^16 (their delays and chirps once, the ring slot of each block's delay: Y_b[k]), and writes
// the Q sub-unit spectra of each block
//   X_c[k'] = e^{+2 pi i c k' / L} sum_s e^{+2 pi i c s / Q} Y_b[k' + s 2^16]   (a Q-point DFT over s)
// at position acc_chirp_pos(perm_pos(k', 2, 14)) of sub-unit c: the order in which acc_unit's transforms
// load them straight into LDS, as they load the chirp. In that order the chunk's bins fill, for each
// sub-unit c and r = k' mod 4, the 16 rows (r, w) of 64 16-byte words (position
// r 2^14 + 1024 w + 128 i + 2 l + e, e = 0 for bin k', 1 for k' + 4096). The workgroup takes one w at a
// time: its 512 threads gather and transform one bin each (threads 256 e + 128 h + m: bin
// k' = 8192 i + 4 base(w) + 512 h + m + 4096 e with base(w) = 32 (w div 4) + 256 (w mod 4), i.e. r = m mod 4,
// word l = m div 4 + 32 h), put the values of 8 / Q blocks at a time (32 rows) into LDS in the rows'
// order, and each wave stores four whole 1 KiB rows per barrier.
// The twiddle e^{+2 pi i k' / L} comes from the plan's split tables (one product of correctly rounded
// factors) and its powers c = 2, 3 from products of it.
// Workgroups take the DMs fastest (neighbouring workgroups gather the same bins of one ring for different
// trials), and the spectra are written with non-temporal stores: they are read back by the next kernel,
// long after any cache would have kept them, and the ring lines stay in the L2s instead.
__device__ __forceinline__ void st_stream(float4 *p, float4 v)
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(__builtin_bit_cast(v4f, v), reinterpret_cast<v4f *>(p));
}

constexpr int PRE_WG = 512;  // threads of a gathering-pass workgroup
constexpr int PRE_CH = 8192; // sub-spectrum bins of a workgroup
constexpr int PRE_CHUNKS = (1 << 16) / PRE_CH;

template <int Q>
__global__ void __launch_bounds__(PRE_WG) acc_pre_kernel(InvArgs a)
{
    static_assert(Q == 2 || Q == 4, "two or four sub-units");
    constexpr int LS = 1 << 16, ROWS = 4 * Q; // rows (c, r) of one w and one block
    constexpr int NBB = 8 / Q;                // blocks per LDS buffer: 32 rows, four per wave
    constexpr int NR = NBB * ROWS;
    constexpr long long L = (long long)Q * LS;
    // two buffers of the rows (j, c, r) of one w for blocks b0 + j: 64 words of 16 bytes each
    __shared__ float4 rows_lds[2][NR][64];
    const int d = (int)(blockIdx.x % a.D), chunk = (int)(blockIdx.x / a.D);
    const int tid = (int)threadIdx.x;
    // the sub-unit launch's arrival counters of this DM's units start at zero
    if (a.sub_sync && chunk == 0)
        for (int b = tid; b < a.B; b += PRE_WG) a.sub_sync[(size_t)d * a.B + b] = 0u;
    const uint16_t *__restrict__ dly = a.delay + (size_t)d * L;
    const float2 *__restrict__ H = a.H + (size_t)d * L;
    float2 *__restrict__ X = a.xpre + (size_t)d * a.B * (size_t)L;
    const int mask = (1 << a.tw_s) - 1;
    const int e = tid >> 8, h = (tid >> 7) & 1, m = tid & 127;
    const int r = m & 3, l = (m >> 2) + 32 * h;
    float *slot = reinterpret_cast<float *>(&rows_lds[0][r][l]) + 2 * e; // this thread's half word, c = 0
    const int wave = tid >> 6, lane = tid & 63;
    int buf = 0;
#pragma unroll 1
    for (int w = 0; w < 16; ++w) {
        const int kp = chunk * PRE_CH + 4 * (32 * (w >> 2) + 256 * (w & 3)) + 512 * h + m + 4096 * e;
        int dd[Q];
        float2 hc[Q];
#pragma unroll
        for (int s = 0; s < Q; ++s) {
            dd[s] = dly[kp + (long long)s * LS];
            hc[s] = H[kp + (long long)s * LS];
        }
        const float2 w1 = cmul(a.tw_hi[kp >> a.tw_s], a.tw_lo[kp & mask]);
        float2 wc[Q];
        wc[0] = make_float2(1.f, 0.f);
#pragma unroll
        for (int c = 1; c < Q; ++c) wc[c] = c == 1 ? w1 : cmul(wc[c - 1], w1);
        // row (c, rr) of this w, word lane: position c 2^16 + rr 2^14 + 1024 w + 128 chunk + 2 lane
        const size_t row0 = (size_t)1024 * w + 128 * chunk + 2 * lane;
        // Q = 2: the ring values of the next NBB blocks are requested before this step's rows are stored
        // (a load issued after the stores would wait for them: they share the wave's memory counter);
        // after the last step the request repeats it (in bounds, unused). Measured at one block per
        // step: 3.78 -> 3.46 ms per launch at Q = 2 (387 trials x 32 blocks), but 4.03 -> 4.33 ms at
        // Q = 4, which gathers each block when it needs it.
        constexpr bool ahead = Q == 2;
        auto gather = [&](int b, float2 *y) {
#pragma unroll
            for (int s = 0; s < Q; ++s) {
                long long sl = (a.g0 + b - (long long)dd[s]) % a.FBS;
                if (sl < 0) sl += a.FBS;
                y[s] = a.ring[sl * L + kp + (long long)s * LS];
            }
        };
        // blocks b0 .. b0 + NBB - 1; indices past the last block repeat it (in bounds; never stored)
        auto gather_step = [&](int b0, float2 (&yy)[NBB][Q]) {
#pragma unroll
            for (int j = 0; j < NBB; ++j) gather(b0 + j < a.B ? b0 + j : a.B - 1, yy[j]);
        };
        float2 yn[NBB][Q];
        if constexpr (ahead) gather_step(0, yn);
#pragma unroll 1
        for (int b0 = 0; b0 < a.B; b0 += NBB) {
            float2 y[NBB][Q];
            if constexpr (ahead) {
#pragma unroll
                for (int j = 0; j < NBB; ++j)
#pragma unroll
                    for (int s = 0; s < Q; ++s) y[j][s] = yn[j][s];
                gather_step(b0 + NBB < a.B ? b0 + NBB : b0, yn);
            } else {
                gather_step(b0, y);
            }
            float *sb = slot + buf * (NR * 64 * 4);
#pragma unroll
            for (int j = 0; j < NBB; ++j) {
#pragma unroll
                for (int s = 0; s < Q; ++s) y[j][s] = cmul(y[j][s], hc[s]);
                dft<+1, Q>(y[j]);
#pragma unroll
                for (int c = 0; c < Q; ++c) {
                    const float2 v = c == 0 ? y[j][0] : cmul(y[j][c], wc[c]);
                    sb[(j * ROWS + c * 4) * 64 * 4] = v.x; // rows_lds[buf][j ROWS + 4 c + r][l], half e
                    sb[(j * ROWS + c * 4) * 64 * 4 + 1] = v.y;
                }
            }
            lds_barrier(); // the rows of this step are in LDS; the other buffer's readers are done
#pragma unroll
            for (int k = 0; k < NR / 8; ++k) {
                const int row = wave + 8 * k; // j ROWS + 4 c + rr
                const int j = row / ROWS, rw = row % ROWS;
                if (b0 + j < a.B) {
                    float2 *Xb = X + (size_t)(b0 + j) * L;
                    st_stream(reinterpret_cast<float4 *>(Xb + (size_t)(rw >> 2) * LS + (size_t)(rw & 3) * (1 << 14) + row0),
                              rows_lds[buf][row][lane]);
                }
            }
            buf ^= 1;
        }
    }
}

// Sub-unit accumulate kernel (L = Q 2^16, plain epilogue: complex64 / uint16 outputs, fused fold) on the
// gathering pass's spectra. Items x = (d B + b) Q + c in the XCD-grouped order: the G/8 workgroups of an
// XCD take consecutive items.
template <int Q, bool RING32, int OUTK>
__global__ void __launch_bounds__(WG) acc_sub_kernel(InvArgs a)
{
    constexpr bool ONE_OUT = OUTK != ACC_OUT_ANY;
    constexpr int R = 4;
    using A = AccShape<R>;
    extern __shared__ float2 lds_all[];
    const int total = __builtin_amdgcn_readfirstlane(a.D * a.B * Q);
    const int G = gridDim.x, w = blockIdx.x;
    // item indices in 32 bits (the launch checks D B Q < 2^31) and in scalar registers: a wide or vector
    // loop counter ends up in a spilled register whose reload would wait for the output stores
    int x = G % 8 == 0 ? (w % 8) * (G / 8) + w / 8 : w;
    auto item = [&](int v, int &d, int &b, int &c) {
        c = __builtin_amdgcn_readfirstlane(v % Q);
        const int u = v / Q;
        b = __builtin_amdgcn_readfirstlane(u % a.B);
        d = __builtin_amdgcn_readfirstlane(u / a.B);
    };
    int d = 0, b = 0, c = 0;
    bool have = x < total; // uniform over the workgroup
    if (have) item(x, d, b, c);
    const AccTw tw = acc_load_twiddles<R, Q>(lds_all, a);
    double *hsum = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_all) + A::LDS_BYTES);
    unsigned int *hcnt = reinterpret_cast<unsigned int *>(hsum + a.nbins);
    float *pw_lds = reinterpret_cast<float *>(hcnt + a.nbins + (a.nbins & 1));
    fold_lds_init(hsum, hcnt, pw_lds, a.fold_sum ? a.nbins : 0);
    uint4 dl = make_uint4(0u, 0u, 0u, 0u); // unused by sub-units
    float2 xv[16]; // one output: the item's first spectrum, loaded by the item before it
    if constexpr (ONE_OUT) {
        if (have) acc_x0_load(a.xpre + (((size_t)d * a.B + b) * Q + c) * (size_t)(1 << 16), xv);
        // wait for them here, so that at the loop head only the previous item's stores follow the loads
#pragma unroll
        for (int t = 0; t < 16; ++t) asm volatile("" : "+v"(xv[t].x), "+v"(xv[t].y));
    }
    // the Q sub-units of a unit are consecutive items: with G a multiple of 8 Q they run in the same round
    // of the walk on one XCD, and wait for each other before writing (otherwise they do not wait)
    const bool sync = a.sub_sync != nullptr && G % (8 * Q) == 0;
#pragma unroll 1
    while (have) {
        const int xn = x + G;
        const bool next = xn < total;
        int dn = 0, bn = 0, cn = 0;
        if (next) item(xn, dn, bn, cn);
        const float2 *xnext = ONE_OUT && next ? a.xpre + (((size_t)dn * a.B + bn) * Q + cn) * (size_t)(1 << 16) : nullptr;
        acc_unit<R, RING32, VAR_PLAIN, Q, OUTK>(a, d, b, c, lds_all, tw, hsum, hcnt, pw_lds, dl, next, dn, bn,
                                                   nullptr, nullptr, sync ? a.sub_sync + (size_t)d * a.B + b : nullptr,
                                                   xnext, xv);
        x = xn;
        d = dn;
        b = bn;
        c = cn;
        have = next;
    }
}

static int acc_r(int log2L) { return log2L == 14 ? 1 : log2L == 15 ? 2 : log2L == 16 || log2L == 17 || log2L == 18 ? 4 : 0; }

int accum_subunits(int log2L) { return log2L == 17 ? 2 : log2L == 18 ? 4 : 1; }

bool accum_supported(int log2L) { return log2L >= 14 && log2L <= 18; }

int accum_lds_bytes(int log2L)
{
    if (!accum_supported(log2L)) return -1;
    switch (acc_r(log2L)) {
    case 1: return AccShape<1>::LDS_BYTES;
    case 2: return AccShape<2>::LDS_BYTES;
    case 4: return AccShape<4>::LDS_BYTES;
    }
    return -1;
}

template <int R>
static hipError_t configure_acc_t(int bytes)
{
    const void *fs[] = {reinterpret_cast<const void *>(&acc_kernel<R, true, VAR_PLAIN>),
                        reinterpret_cast<const void *>(&acc_kernel<R, true, VAR_F16>),
                        reinterpret_cast<const void *>(&acc_kernel<R, true, VAR_TANGENT>),
                        reinterpret_cast<const void *>(&acc_kernel<R, false, VAR_PLAIN>),
                        reinterpret_cast<const void *>(&acc_kernel<R, false, VAR_F16>),
                        reinterpret_cast<const void *>(&acc_kernel<R, false, VAR_TANGENT>)};
    for (const void *f : fs) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int Q>
static hipError_t configure_acc_sub_t(int bytes)
{
    const void *fs[] = {reinterpret_cast<const void *>(&acc_sub_kernel<Q, true, ACC_OUT_ANY>),
                        reinterpret_cast<const void *>(&acc_sub_kernel<Q, true, ACC_OUT_C64>),
                        reinterpret_cast<const void *>(&acc_sub_kernel<Q, true, ACC_OUT_U16>)};
    for (const void *f : fs) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t configure_accum_lds(int log2L, int bytes)
{
    if (log2L == 17) return configure_acc_sub_t<2>(bytes);
    if (log2L == 18) return configure_acc_sub_t<4>(bytes);
    switch (log2L) {
    case 14: return configure_acc_t<1>(bytes);
    case 15: return configure_acc_t<2>(bytes);
    case 16: return configure_acc_t<4>(bytes);
    }
    return hipErrorInvalidValue;
}

template <int R>
static hipError_t launch_acc_t(const InvArgs &a, int grid, hipStream_t st)
{
    using A = AccShape<R>;
    const size_t lds = A::LDS_BYTES + (a.fold_sum ? (size_t)a.nbins * 12 : 0) + INV_EXTRA_LDS;
    const bool ring32 = (unsigned long long)a.FBS * A::L * sizeof(float2) < (1ull << 32);
    const int var = a.tangent_ref ? VAR_TANGENT : a.out_f16 ? VAR_F16 : VAR_PLAIN;
#define LAUNCH(RG, V) hipLaunchKernelGGL((acc_kernel<R, RG, V>), dim3(grid), dim3(WG), lds, st, a)
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

template <int Q>
static hipError_t launch_acc_sub_t(const InvArgs &a, int grid, hipStream_t st)
{
    using A = AccShape<4>;
    // the sub-unit form has the plain epilogue only; the caller chooses another form otherwise
    if (a.tangent_ref || a.out_f16 || !a.sub_lo || !a.xpre) return hipErrorInvalidValue;
    if ((long long)a.D * a.B * Q >= (1ll << 31)) return hipErrorInvalidValue; // 32-bit item indices
    const unsigned pre_grid = (unsigned)((long long)a.D * PRE_CHUNKS);
    hipLaunchKernelGGL((acc_pre_kernel<Q>), dim3(pre_grid), dim3(PRE_WG), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = A::LDS_BYTES + (a.fold_sum ? (size_t)a.nbins * 12 : 0) + INV_EXTRA_LDS;
    // exactly one output and no fold: the sub-units load each next spectrum ahead of their stores
    if (!a.fold_sum && a.out_c64 && !a.out_u16)
        hipLaunchKernelGGL((acc_sub_kernel<Q, true, ACC_OUT_C64>), dim3(grid), dim3(WG), lds, st, a);
    else if (!a.fold_sum && !a.out_c64 && a.out_u16)
        hipLaunchKernelGGL((acc_sub_kernel<Q, true, ACC_OUT_U16>), dim3(grid), dim3(WG), lds, st, a);
    else
        hipLaunchKernelGGL((acc_sub_kernel<Q, true, ACC_OUT_ANY>), dim3(grid), dim3(WG), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_accum(int log2L, const InvArgs &a, int grid, hipStream_t st)
{
    switch (log2L) {
    case 14: return launch_acc_t<1>(a, grid, st);
    case 15: return launch_acc_t<2>(a, grid, st);
    case 16: return launch_acc_t<4>(a, grid, st);
    case 17: return launch_acc_sub_t<2>(a, grid, st);
    case 18: return launch_acc_sub_t<4>(a, grid, st);
    }
    return hipErrorInvalidValue;
}

} // namespace msosm
