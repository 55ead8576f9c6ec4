// One persistent dedispersion launch over DM trials of different FFT lengths (uneven DM grids).
// Each workgroup takes (class, DM, block) units from a device counter in increasing order (the
// classes are sorted by decreasing length, so the longest transforms are scheduled first and the
// short ones fill the tail) and runs the unit with the same device code as the one-length kernel.
#include "msosm_device.h"

namespace msosm {

// lengths 2^14 .. 2^21: the range of per-DM lengths of uneven grids
#define MSOSM_MULTI_SHAPES(X)                                                                                          \
    X(14, 128, 128) X(15, 256, 128) X(16, 256, 256) X(17, 512, 256) X(18, 512, 512) X(19, 1024, 512)                  \
    X(20, 1024, 1024) X(21, 2048, 1024)

// the multi-length kernel reads every class's twiddle tables from global memory: tiles only
static constexpr int multi_tiles_bytes()
{
    int m = 0;
#define X(lg, a, b) m = m > Shape<a, b>::LDS_TILES_BYTES ? m : Shape<a, b>::LDS_TILES_BYTES;
    MSOSM_MULTI_SHAPES(X)
#undef X
    return m;
}

template <int VAR>
__global__ void __launch_bounds__(WG) inv_multi_kernel(MultiInvArgs a)
{
    extern __shared__ float2 lds_all[];
    constexpr int HOFF = multi_tiles_bytes();
    double *hsum = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_all) + HOFF);
    unsigned int *hcnt = reinterpret_cast<unsigned int *>(hsum + a.nbins);
    float *pw_lds = reinterpret_cast<float *>(hcnt + a.nbins + (a.nbins & 1));
    int *ticket = reinterpret_cast<int *>(pw_lds + 1);
    fold_lds_init(hsum, hcnt, pw_lds, a.nbins);
    float2 *scr = a.scratch + (size_t)blockIdx.x * a.scr_stride;
    const long long U = a.ustart[a.ncls];
    int c = 0;
#pragma unroll 1
    for (;;) {
        if (threadIdx.x == 0) *ticket = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        // the ticket is the same in every lane: make it wave-uniform so that the class switch and
        // the unit's scalars stay in scalar registers (a divergent switch would merge the register
        // demands of all lengths)
        const long long u = __builtin_amdgcn_readfirstlane(*ticket);
        __syncthreads(); // every lane has read the ticket before the next one is written
        if (u >= U) break;
        while (u >= a.ustart[c + 1]) ++c; // a workgroup's tickets increase, so c never decreases
        const InvArgs &ca = a.cls[c];
        const long long lu = u - a.ustart[c];
        const int d = (int)(lu / ca.B), b = (int)(lu % ca.B);
        const Tw tw{ca.tw_n1, ca.tw_n2, ca.tw_lo, ca.tw_hi, ca.tw_s};
        switch (a.log2L[c]) {
#define X(lg, n1, n2)                                                                                                  \
    case lg:                                                                                                           \
        inv_unit<n1, n2, true, VAR, VAR == VAR_PLAIN && lg <= 16, SUBT, VAR == VAR_PLAIN>(ca, d, b, lds_all, tw, scr,  \
                                                                                         hsum, hcnt, pw_lds);          \
        break;
            MSOSM_MULTI_SHAPES(X)
#undef X
        }
    }
}

bool multi_length_supported(int log2L) { return log2L >= 14 && log2L <= 21; }

int multi_tiles_lds_bytes() { return multi_tiles_bytes(); }

hipError_t configure_multi_lds(int bytes)
{
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&inv_multi_kernel<VAR_PLAIN>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(&inv_multi_kernel<VAR_F16>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

NodeDesc inv_multi_node(const MultiInvArgs &a, int grid)
{
    NodeDesc n;
    n.func = a.cls[0].out_f16 ? reinterpret_cast<const void *>(&inv_multi_kernel<VAR_F16>)
                              : reinterpret_cast<const void *>(&inv_multi_kernel<VAR_PLAIN>);
    n.grid = dim3(grid);
    n.block = dim3(WG);
    n.lds = (unsigned int)(multi_tiles_bytes() + (size_t)a.nbins * 12 + INV_EXTRA_LDS);
    return n;
}

hipError_t launch_inv_multi(const MultiInvArgs &a, int grid, hipStream_t st)
{
    const size_t lds = (size_t)multi_tiles_bytes() + (size_t)a.nbins * 12 + INV_EXTRA_LDS;
    if (a.cls[0].out_f16)
        hipLaunchKernelGGL(inv_multi_kernel<VAR_F16>, dim3(grid), dim3(WG), lds, st, a);
    else
        hipLaunchKernelGGL(inv_multi_kernel<VAR_PLAIN>, dim3(grid), dim3(WG), lds, st, a);
    return hipGetLastError();
}

} // namespace msosm
