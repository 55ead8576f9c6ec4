// Persistent dedispersion kernel of the accumulate form (FFT lengths 2^14, 2^15, 2^16): one workgroup
// per CU walks (DM, block) units in the plan's unit order; each unit's inverse transform stays in the
// CU (LDS and registers), so the launch reads the gathered spectra and writes only its outputs.
#include "msosm_accum.h"

namespace msosm {

template <int R, bool RING32, int VAR>
__global__ void __launch_bounds__(WG) acc_kernel(InvArgs a)
{
    using A = AccShape<R>;
    extern __shared__ float2 lds_all[];
    uint32_t *slots = reinterpret_cast<uint32_t *>(lds_all + ACC_AREA + A::TW_WORDS); // two tables of 256
    const UnitOrder o(a, gridDim.x, blockIdx.x);
    long long s = o.first;
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
#pragma unroll 1
    while (have) {
        int dn = 0, bn = 0;
        const bool next = o.next(s, a, dn, bn);
        acc_unit<R, RING32, VAR>(a, d, b, lds_all, tw, hsum, hcnt, pw_lds, dl, next, dn, bn, slots + 256 * cur,
                                 slots + 256 * (cur ^ 1));
        cur ^= 1;
        d = dn;
        b = bn;
        have = next;
    }
}

static int acc_r(int log2L) { return log2L == 14 ? 1 : log2L == 15 ? 2 : log2L == 16 ? 4 : 0; }

bool accum_supported(int log2L) { return acc_r(log2L) != 0; }

int accum_lds_bytes(int log2L)
{
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

hipError_t configure_accum_lds(int log2L, int bytes)
{
    switch (acc_r(log2L)) {
    case 1: return configure_acc_t<1>(bytes);
    case 2: return configure_acc_t<2>(bytes);
    case 4: return configure_acc_t<4>(bytes);
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

hipError_t launch_accum(int log2L, const InvArgs &a, int grid, hipStream_t st)
{
    switch (acc_r(log2L)) {
    case 1: return launch_acc_t<1>(a, grid, st);
    case 2: return launch_acc_t<2>(a, grid, st);
    case 4: return launch_acc_t<4>(a, grid, st);
    }
    return hipErrorInvalidValue;
}

} // namespace msosm
