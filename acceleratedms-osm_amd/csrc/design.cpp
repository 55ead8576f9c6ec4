// Host-side design of the MS-OSM path: dispersion smear Nd, filter order M, segmentation into
// delay blocks, segment runs and the dedispersion transfer function H.
//
// This translation unit must be compiled with -ffp-contract=off: Nd, delaycount and the segment
// boundaries are integer results of mixed float/double expressions, and they match the reference
// host build only when every operation is evaluated as written (x86-64 SSE, no FMA contraction).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "capi_internal.h"
#include "msosm_amd.h"

namespace msosm {

namespace {

constexpr double kPi = 3.14159265358979323846; // utils/include/globals.h:30 (M_PI)

// FP64 FFT (iterative radix-2) used only for the transfer-function design.
class Fft64 {
public:
    explicit Fft64(long n) : n_(n), tw_(n / 2)
    {
        for (long k = 0; k < n / 2; ++k) tw_[k] = unit_root(k, n);
    }
    // sign -1: e^{-i}, +1: e^{+i}; unnormalised, in place
    void run(std::complex<double> *x, int sign) const
    {
        const long n = n_;
        for (long i = 1, j = 0; i < n; ++i) {
            long bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(x[i], x[j]);
        }
        for (long len = 2; len <= n; len <<= 1) {
            const long half = len >> 1, stride = n / len;
            for (long i = 0; i < n; i += len)
                for (long j = 0; j < half; ++j) {
                    std::complex<double> w = tw_[j * stride];
                    if (sign < 0) w = std::conj(w);
                    const std::complex<double> b = x[i + j + half] * w;
                    x[i + j + half] = x[i + j] - b;
                    x[i + j] += b;
                }
        }
    }
    // e^{+2 pi i k/n} accurate to ~1 ulp via reduction to the first octant
    static std::complex<double> unit_root(long k, long n)
    {
        long kk = ((k % n) + n) % n;
        bool neg_s = false, neg_c = false, swp = false;
        if (2 * kk > n) { kk = n - kk; neg_s = true; }
        if (4 * kk > n) { kk = n / 2 - kk; neg_c = true; }
        if (8 * kk > n) { kk = n / 4 - kk; swp = true; }
        const double a = 2.0 * kPi * (double)kk / (double)n;
        double c = std::cos(a), s = std::sin(a);
        if (swp) std::swap(c, s);
        if (neg_c) c = -c;
        if (neg_s) s = -s;
        return {c, s};
    }

private:
    long n_;
    std::vector<std::complex<double>> tw_;
};

} // namespace

int design_nd(float bw, float f0, float dm)
{
    // msosm_prepare.cpp:27-28,33 with float members fs, kdm (msosm_prepare.h:35-39)
    const float fs = bw;
    const float kdm = (float)(4.15 * 1e15);
    const double inner = 1 / std::pow((double)f0, 2) - 1 / std::pow((double)(float)(f0 + bw), 2);
    const float inv_fs = 1 / fs;
    const double v = (double)(float)(kdm * dm) * inner / (double)inv_fs;
    return (int)std::floor(v);
}

int design_min_order(int nd)
{
    // msosm_prepare.cpp:59-61, MIN_SEGMENT_POINTS = 1024 (msosm_prepare.h:7)
    const double order = std::sqrt(1024 / 2.0) * std::sqrt((double)nd);
    return (int)std::pow(2, std::ceil(std::log2(order)));
}

unsigned long design_osm_order(unsigned long nd)
{
    // osm_prepare.cpp:21-23 and next_power_of_2 (osm_prepare.cpp:127-140)
    int n = (int)nd;
    int p;
    if (n <= 0) {
        p = 1;
    } else {
        --n;
        n |= n >> 1;
        n |= n >> 2;
        n |= n >> 4;
        n |= n >> 8;
        n |= n >> 16;
        p = n + 1;
    }
    return 2ul * (unsigned long)p;
}

int design_delaycount(int nd, int M)
{
    // msosm_prepare.cpp:217; Nd == 0 defined as one block (no delay)
    if (nd <= 0) return 1;
    const float q = (float)nd / (float)M;
    return (int)std::ceil(q);
}

int design_segmentation(float bw, float f0, float dm, int M, int *delay_points)
{
    // msosm_prepare.cpp:215-284, using this DM's own delaycount for the positive-side count
    const float fs = bw;
    const float kdm = (float)(4.15 * 1e15);
    const float w0 = (float)(2 * kPi * (double)f0);
    const int nd = design_nd(bw, f0, dm);
    const int dc = design_delaycount(nd, M);
    if (nd <= 0) {
        std::fill(delay_points, delay_points + 2 * (size_t)M, 0);
        return 1;
    }
    std::vector<float> w(dc + 1);
    w[0] = -1;
    w[dc] = 1;
    int pos_index = 1;
    const double denom = 4 * kPi * kPi * (double)kdm * (double)dm;
    for (int j = 1; j < dc; ++j) {
        const int jm = j * M;
        const float jm_fs = (float)jm / fs;
        const double t = 1 / std::pow((double)w0, 2) - (double)jm_fs / denom;
        w[j] = (float)((std::sqrt(1 / t) - (double)w0 - kPi * (double)fs) / kPi / (double)fs);
        if (w[j] < 0) pos_index = j + 1;
    }
    const int negcount = pos_index;
    const int poscount = dc - pos_index + 1;
    std::vector<int> segneg(negcount + 1), segpos(poscount + 1);
    for (int j = 0; j < negcount; ++j) segneg[j] = (int)std::floor((w[j] + (float)2) * (float)M);
    segneg[negcount] = 2 * M;
    segpos[0] = 0;
    for (int j = 0; j < poscount; ++j) segpos[j + 1] = (int)std::floor(w[j + negcount] * (float)M);
    std::fill(delay_points, delay_points + 2 * (size_t)M, -1);
    for (int j = 0; j < negcount; ++j)
        for (int k = segneg[j]; k < segneg[j + 1]; ++k) delay_points[k] = j;
    for (int j = 0; j < poscount; ++j)
        for (int k = segpos[j]; k < segpos[j + 1]; ++k) delay_points[k] = negcount - 1 + j;
    return dc;
}

int design_build_segments(const int *delay, int fft_len, int max_len, int *seg, int cap)
{
    // msosm_gpu_batch.cpp:262-295
    if (fft_len <= 0) return 0;
    int count = 0;
    int cur_group = delay[0], cur_start = 0;
    auto emit = [&](int start, int len, int group) -> bool {
        if (count >= cap) return false;
        seg[3 * count] = start;
        seg[3 * count + 1] = len;
        seg[3 * count + 2] = group;
        ++count;
        return true;
    };
    for (int k = 1; k < fft_len; ++k) {
        if (delay[k] != cur_group || k - cur_start == max_len) {
            if (!emit(cur_start, k - cur_start, cur_group)) return -1;
            cur_start = k;
            cur_group = delay[k];
        }
    }
    if (!emit(cur_start, fft_len - cur_start, cur_group)) return -1;
    return count;
}

// Chirp sample n of the M-point design grid: c_n = e^{-i phi(f_n)}, f_n = ((n + M/2) mod M) fs/M
// above the band bottom (msosm_prepare.cpp:122-138; legacy float-w form at :179-192).
static double chirp_phase(float bw, float f0, float dm, int M, int n, int phase_mode)
{
    const int j = (n < M / 2) ? n + M / 2 : n - M / 2; // grid index of bin n
    double phase;
    if (phase_mode == MSOSM_PHASE_FP64) {
        const double kdm_d = 4.15e15, fs_d = (double)bw, dm_d = (double)dm, f0_d = (double)f0;
        const double w0_d = 2.0 * kPi * f0_d;
        const double step = 2.0 * kPi * fs_d / (double)M;
        const double w = step * (double)j;
        phase = 4.0 * kPi * kPi * kdm_d * dm_d * w * w / ((w + w0_d) * w0_d * w0_d);
    } else {
        const float fs = bw, kdm = (float)(4.15 * 1e15), w0 = (float)(2 * kPi * (double)f0);
        const float step = (float)(2 * kPi * (double)fs / (double)M);
        float w = (float)(-kPi * (double)fs + (double)(step * (float)j));
        w = (float)((double)w + kPi * (double)fs);
        phase = 4 * kPi * kPi * (double)kdm * (double)dm * (double)w * (double)w / (double)(w + w0) / (double)w0 /
                (double)w0;
    }
    return phase;
}

static std::complex<double> chirp(float bw, float f0, float dm, int M, int n, int phase_mode)
{
    const double phase = chirp_phase(bw, f0, dm, M, n, phase_mode);
    return {std::cos(phase), -std::sin(phase)};
}

// d c_n / d DM = -i (phi_n / DM) c_n: the phase is linear in DM, so phi_n / DM is the phase at DM = 1
static std::complex<double> chirp_ddm(float bw, float f0, float dm, int M, int n, int phase_mode)
{
    const double unit = chirp_phase(bw, f0, 1.0f, M, n, phase_mode);
    const std::complex<double> c = chirp(bw, f0, dm, M, n, phase_mode);
    return std::complex<double>(0.0, -unit) * c;
}

// H = DFT_2M([IDFT_M(c)/M, 0]) (msosm_prepare.cpp:140-161), evaluated in FP64 as
//   H[2n] = c_n, H[2n+1] = DFT_M(e^{-i pi j/M} IDFT_M(c)[j] / M)[n],
// then rounded once to complex64.
// With deriv != 0 the same map is applied to d c / d DM, which gives d H / d DM (the map is linear).
int design_transfer_function(float bw, float f0, float dm, int M, int phase_mode, float *H, const Fft64Handle *fft,
                             int deriv)
{
    std::vector<std::complex<double>> c(M), h(M);
    for (int n = 0; n < M; ++n)
        c[n] = deriv ? chirp_ddm(bw, f0, dm, M, n, phase_mode) : chirp(bw, f0, dm, M, n, phase_mode);
    h = c;
    const Fft64 *f = reinterpret_cast<const Fft64 *>(fft);
    Fft64 local(fft ? 1 : M);
    if (!f) f = &local;
    f->run(h.data(), +1);
    for (int j = 0; j < M; ++j) h[j] *= Fft64::unit_root(-(long)j, 2L * M) / (double)M;
    f->run(h.data(), -1);
    for (int n = 0; n < M; ++n) {
        H[4 * (size_t)n] = (float)c[n].real();
        H[4 * (size_t)n + 1] = (float)c[n].imag();
        H[4 * (size_t)n + 2] = (float)h[n].real();
        H[4 * (size_t)n + 3] = (float)h[n].imag();
    }
    return 0;
}

void unit_root(long k, long n, double *re, double *im)
{
    const std::complex<double> w = Fft64::unit_root(k, n);
    *re = w.real();
    *im = w.imag();
}

const Fft64Handle *design_fft_create(int M) { return reinterpret_cast<const Fft64Handle *>(new Fft64(M)); }
void design_fft_destroy(const Fft64Handle *h) { delete reinterpret_cast<const Fft64 *>(h); }
void design_fft_run(const Fft64Handle *h, std::complex<double> *x, int sign)
{
    reinterpret_cast<const Fft64 *>(h)->run(x, sign);
}

int design_transfer_functions(float bw, float f0, const float *dms, int num_dms, int M, int phase_mode, float *H,
                              int nthreads, int deriv)
{
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min(nthreads, num_dms);
    const Fft64Handle *fft = design_fft_create(M);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < num_dms; i += nthreads)
                design_transfer_function(bw, f0, dms[i], M, phase_mode, H + 4 * (size_t)M * i, fft, deriv);
        });
    for (auto &th : pool) th.join();
    design_fft_destroy(fft);
    return 0;
}

} // namespace msosm
