/*
 * CPU oracle for the MS-OSM coherent-dedispersion path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Plain C, FP64 restatement of the reference algorithm. Build with
 *   -O2 -ffp-contract=off -fno-fast-math
 * so that the design and fold arithmetic (float/double mixes copied in expression order from the
 * reference) is evaluated exactly as the reference's x86-64 host build evaluates it.
 * Paths in citations are relative to the reference repository root.
 */
#include "msosm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------------------------------
 * FFT (FP64). Iterative radix-2 DIT with an accurately generated twiddle table (octant symmetry,
 * so every twiddle is within ~1 ulp of e^{2 pi i k/n}). Not used by the product.
 * --------------------------------------------------------------------------------------------- */
static void twiddle_exact(long k, long n, double *c, double *s)
{
    /* e^{+2 pi i k/n} for 0 <= k < n, reduced to the first octant */
    /* reduce k into [0, n/8] using symmetries of cos/sin */
    long kk = k % n;
    int neg_s = 0, neg_c = 0, swap = 0;
    if (kk > n / 2) { kk = n - kk; neg_s = 1; }          /* angle -> 2pi - angle */
    if (kk > n / 4) { kk = n / 2 - kk; neg_c = 1; }      /* angle -> pi - angle */
    if (kk > n / 8) { kk = n / 4 - kk; swap = 1; }       /* angle -> pi/2 - angle */
    double a = 2.0 * M_PI * (double)kk / (double)n;
    double cc = cos(a), ss = sin(a);
    if (swap) { double t = cc; cc = ss; ss = t; }
    if (neg_c) cc = -cc;
    if (neg_s) ss = -ss;
    *c = cc;
    *s = ss;
}

int orc_fft(double *x, long n, int sign)
{
    if (n < 1 || (n & (n - 1)) != 0) return -1;
    if (n == 1) return 0;
    /* bit reversal */
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j];
            x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr;
            x[2 * j + 1] = ti;
        }
    }
    double *tw = (double *)malloc(sizeof(double) * (size_t)n); /* n/2 complex */
    if (!tw) return -2;
    for (long k = 0; k < n / 2; k++) {
        double c, s;
        twiddle_exact(k, n, &c, &s);
        tw[2 * k] = c;
        tw[2 * k + 1] = sign < 0 ? -s : s;
    }
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1, step = n / len;
        for (long i = 0; i < n; i += len) {
            for (long j = 0; j < half; j++) {
                double wr = tw[2 * j * step], wi = tw[2 * j * step + 1];
                double *a = x + 2 * (i + j), *b = x + 2 * (i + j + half);
                double br = b[0] * wr - b[1] * wi;
                double bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br;
                b[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
    free(tw);
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Design. Types and expression order follow the reference so the results are bit-identical:
 * members bw, fs, kdm, f0, w0 are float (dedispersion/include/msosm_prepare.h:33-43).
 * --------------------------------------------------------------------------------------------- */
static const double ref_pi = M_PI; /* utils/include/globals.h:30 (constexpr auto pi = M_PI) */

int orc_nd(float bw, float f0, float dm)
{
    /* msosm_prepare.cpp:5-7,27-28,33 */
    float fs = bw;
    float kdm = (float)(4.15 * 1e15);
    double inner = 1 / pow((double)f0, 2) - 1 / pow((double)(float)(f0 + bw), 2);
    float inv_fs = 1 / fs; /* int / float -> float */
    double v = (double)(float)(kdm * dm) * inner / (double)inv_fs;
    return (int)floor(v);
}

int orc_min_order(int nd)
{
    /* msosm_prepare.cpp:59-61 with MIN_SEGMENT_POINTS = 1024 (msosm_prepare.h:7) */
    double order = sqrt(1024 / 2.0) * sqrt((double)nd);
    return (int)pow(2, ceil(log2(order)));
}

long orc_osm_min_order(unsigned long nd)
{
    /* osm_prepare.cpp:17-30: M = 2 * next_power_of_2((int)Nd); next_power_of_2 at :127-140 */
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
    return 2L * p;
}

int orc_delaycount(int nd, int M)
{
    /* msosm_prepare.cpp:217: ceil(Nd / (float)M) evaluated in float */
    if (M <= 0) return -1;
    if (nd <= 0) return 1;
    float q = (float)nd / (float)M;
    return (int)ceilf(q);
}

int orc_segmentation(float bw, float f0, float dm, int M, int *delay_points)
{
    /* msosm_prepare.cpp:215-284 */
    if (M <= 0 || !delay_points) return -1;
    const float fs = bw;
    const float kdm = (float)(4.15 * 1e15);
    const float w0 = (float)(2 * ref_pi * (double)f0);
    const int nd = orc_nd(bw, f0, dm);
    const int dc = orc_delaycount(nd, M);
    if (nd <= 0) {
        /* DM = 0 (or Nd = 0): the reference leaves the table uninitialised; defined as no delay */
        for (int k = 0; k < 2 * M; k++) delay_points[k] = 0;
        return 1;
    }
    float *w = (float *)malloc(sizeof(float) * (size_t)(dc + 1));
    if (!w) return -1;
    w[0] = -1;
    w[dc] = 1;
    int pos_index = 1;
    const double denom = 4 * ref_pi * ref_pi * (double)kdm * (double)dm;
    for (int j = 1; j < dc; j++) {
        /* (sqrt(1 / (1 / pow(w0, 2) - j * M / fs / (4*pi*pi*kdm*dm))) - w0 - pi * fs) / pi / fs */
        int jm = j * M;
        float jm_fs = (float)jm / fs;
        double t = 1 / pow((double)w0, 2) - (double)jm_fs / denom;
        double v = (sqrt(1 / t) - (double)w0 - ref_pi * (double)fs) / ref_pi / (double)fs;
        w[j] = (float)v;
        if (w[j] < 0) pos_index = j + 1;
    }
    const int negcount = pos_index;
    const int poscount = dc - pos_index + 1;
    int *segneg = (int *)malloc(sizeof(int) * (size_t)(negcount + 1));
    int *segpos = (int *)malloc(sizeof(int) * (size_t)(poscount + 1));
    if (!segneg || !segpos) {
        free(w);
        free(segneg);
        free(segpos);
        return -1;
    }
    for (int j = 0; j < negcount; j++) {
        float a = w[j] + (float)2;
        float b = a * (float)M;
        segneg[j] = (int)floorf(b);
    }
    segneg[negcount] = 2 * M;
    segpos[0] = 0;
    for (int j = 0; j < poscount; j++) {
        float b = w[j + negcount] * (float)M;
        segpos[j + 1] = (int)floorf(b);
    }
    for (int k = 0; k < 2 * M; k++) delay_points[k] = -1;
    for (int j = 0; j < negcount; j++)
        for (int k = segneg[j]; k < segneg[j + 1]; k++) delay_points[k] = j;
    for (int j = 0; j < poscount; j++)
        for (int k = segpos[j]; k < segpos[j + 1]; k++) delay_points[k] = negcount - 1 + j;
    free(w);
    free(segneg);
    free(segpos);
    return dc;
}

int orc_build_segments(const int *delay, int fft_len, int max_len, int *seg, int cap)
{
    /* msosm_gpu_batch.cpp:262-295 */
    if (fft_len <= 0) return 0;
    int count = 0;
    int cur_group = delay[0];
    int cur_start = 0;
    for (int k = 1; k < fft_len; ++k) {
        if (delay[k] != cur_group || k - cur_start == max_len) {
            if (count >= cap) return -1;
            seg[3 * count] = cur_start;
            seg[3 * count + 1] = k - cur_start;
            seg[3 * count + 2] = cur_group;
            count++;
            cur_start = k;
            cur_group = delay[k];
        }
    }
    if (count >= cap) return -1;
    seg[3 * count] = cur_start;
    seg[3 * count + 1] = fft_len - cur_start;
    seg[3 * count + 2] = cur_group;
    return count + 1;
}

static int transfer_impl(float bw, float f0, float dm, int M, int phase_fp32, int deriv, double *H);

int orc_transfer_function(float bw, float f0, float dm, int M, int phase_fp32, double *H)
{
    return transfer_impl(bw, f0, dm, M, phase_fp32, 0, H);
}

int orc_transfer_function_ddm(float bw, float f0, float dm, int M, int phase_fp32, double *dH)
{
    return transfer_impl(bw, f0, dm, M, phase_fp32, 1, dH);
}

/* phase of design bin j at a given DM (msosm_prepare.cpp:122-134; legacy :179-192) */
static double chirp_phase(float bw, float f0, float dm, int M, int j, int phase_fp32)
{
    /* msosm_prepare.cpp:107-164 (FP64 phase) and the legacy FP32 phase of :166-213 */
    const double pi_d = 3.14159265358979323846;
    const double kdm_d = 4.15e15;
    const double fs_d = (double)bw;
    const double dm_d = (double)dm;
    const double f0_d = (double)f0;
    const double w0_d = 2.0 * pi_d * f0_d;
    const double step = 2.0 * pi_d * fs_d / (double)M;
    {
        double phase;
        if (!phase_fp32) {
            double w = step * (double)j;
            phase = 4.0 * pi_d * pi_d * kdm_d * dm_d * w * w / ((w + w0_d) * w0_d * w0_d);
        } else {
            /* legacy form: float w[j]; float-valued members; double pi; evaluated as written at
             * msosm_prepare.cpp:181-183 with kdm, w0, fs float members */
            const float fs = bw;
            const float kdm = (float)(4.15 * 1e15);
            const float w0 = (float)(2 * ref_pi * (double)f0);
            float stepf = (float)(2 * ref_pi * (double)fs / (double)M);
            float wf = (float)(-ref_pi * (double)fs + (double)(stepf * (float)j));
            wf = (float)((double)wf + ref_pi * (double)fs);
            double num = 4 * ref_pi * ref_pi * (double)kdm * (double)dm * (double)wf * (double)wf;
            phase = num / (double)(wf + w0) / (double)w0 / (double)w0;
        }
        return phase;
    }
}

/* H = DFT_2M([IDFT_M(c)/M, 0]) with c_n = e^{-i phase}; with deriv, c is replaced by
 * dc/dDM = -i (phase at DM 1) c (the phase is proportional to DM), which gives dH/dDM */
static int transfer_impl(float bw, float f0, float dm, int M, int phase_fp32, int deriv, double *H)
{
    if (M <= 0 || (M & (M - 1)) != 0 || !H) return -1;
    const long L = 2L * M;
    double *h = (double *)calloc((size_t)(2 * M), sizeof(double));
    if (!h) return -1;
    for (int j = 0; j < M; j++) {
        const double phase = chirp_phase(bw, f0, dm, M, j, phase_fp32);
        const double cr = cos(phase), ci = -sin(phase);
        int n = (j < M / 2) ? (M / 2 + j) : (j - M / 2);
        if (!deriv) {
            h[2 * n] = cr;
            h[2 * n + 1] = ci;
        } else {
            const double u = chirp_phase(bw, f0, 1.0f, M, j, phase_fp32);
            h[2 * n] = u * ci; /* -i u (cr + i ci) = u ci - i u cr */
            h[2 * n + 1] = -u * cr;
        }
    }
    /* h = IDFT_M(c) / M (e^{+i}); H = DFT_2M([h, 0]) (e^{-i}) */
    if (orc_fft(h, M, +1) != 0) {
        free(h);
        return -1;
    }
    memset(H, 0, sizeof(double) * (size_t)(2 * L));
    for (int j = 0; j < M; j++) {
        H[2 * j] = h[2 * j] / (double)M;
        H[2 * j + 1] = h[2 * j + 1] / (double)M;
    }
    free(h);
    return orc_fft(H, L, -1);
}

/* ------------------------------------------------------------------------------------------------
 * MS-OSM (FP64), whole record at once; see the header for the definition.
 * --------------------------------------------------------------------------------------------- */
int orc_msosm(const double *x, long n_samples, int M, int num_dms, const int *delays, const double *H,
              double *out)
{
    if (M <= 0 || n_samples % M != 0 || num_dms <= 0) return -1;
    const long L = 2L * M;
    const long nb = n_samples / M;
    double *spec = (double *)malloc(sizeof(double) * (size_t)(2 * L * nb));
    double *y = (double *)malloc(sizeof(double) * (size_t)(2 * L));
    if (!spec || !y) {
        free(spec);
        free(y);
        return -2;
    }
    for (long g = 0; g < nb; g++) {
        double *X = spec + 2 * L * g;
        for (long i = 0; i < L; i++) {
            long s = (g - 1) * M + i;
            X[2 * i] = s < 0 ? 0.0 : x[2 * s];
            X[2 * i + 1] = s < 0 ? 0.0 : x[2 * s + 1];
        }
        orc_fft(X, L, -1);
    }
    for (int d = 0; d < num_dms; d++) {
        const int *dl = delays + (long)d * L;
        const double *Hd = H + 2 * (long)d * L;
        double *od = out + 2 * (long)d * n_samples;
        for (long g = 0; g < nb; g++) {
            for (long k = 0; k < L; k++) {
                long src = g - dl[k];
                if (src < 0) {
                    y[2 * k] = 0.0;
                    y[2 * k + 1] = 0.0;
                } else {
                    const double *X = spec + 2 * L * src;
                    double xr = X[2 * k], xi = X[2 * k + 1];
                    double hr = Hd[2 * k], hi = Hd[2 * k + 1];
                    y[2 * k] = xr * hr - xi * hi;
                    y[2 * k + 1] = xr * hi + xi * hr;
                }
            }
            orc_fft(y, L, +1);
            const double sc = 1.0 / (double)L;
            for (long m = 0; m < M; m++) {
                od[2 * (g * M + m)] = y[2 * (M + m)] * sc;
                od[2 * (g * M + m) + 1] = y[2 * (M + m) + 1] * sc;
            }
        }
    }
    free(spec);
    free(y);
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Fold: gpu/gpu_kernel.cu:462-485, folding/fold_gpu.cpp:48-54.
 * --------------------------------------------------------------------------------------------- */
long orc_fold_bin(long i, unsigned long nbins, double t_start, double dt, double period)
{
    double t = t_start + (double)i * dt;
    double phase = fmod(t / period, 1.0);
    unsigned long bin = (unsigned long)floor(phase * (double)nbins);
    if (bin >= nbins) bin = nbins - 1;
    return (long)bin;
}

void orc_fold(const double *intensity, long size, unsigned long nbins, double t_start, double dt,
              double period, double *sums, unsigned long long *counts, long *bins_out)
{
    for (long i = 0; i < size; i++) {
        double v = intensity[i];
        if (v == 0.0) {
            if (bins_out) bins_out[i] = -1;
            continue;
        }
        long b = orc_fold_bin(i, nbins, t_start, dt, period);
        if (bins_out) bins_out[i] = b;
        sums[b] += v;
        counts[b] += 1ULL;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Codecs.
 * --------------------------------------------------------------------------------------------- */
void orc_unpack_u16(const uint32_t *packed, long n, double *out)
{
    /* gpu_kernel.cu:22-29 */
    for (long i = 0; i < n; i++) {
        uint32_t v = packed[i];
        uint16_t x = (uint16_t)(v & 0xFFFFu), y = (uint16_t)(v >> 16);
        float fx = (float)x - 32768.0f * (float)(x != 0);
        float fy = (float)y - 32768.0f * (float)(y != 0);
        out[2 * i] = fx;
        out[2 * i + 1] = fy;
    }
}

static uint32_t enc16(float x)
{
    /* __float2uint_rn(x + (x != 0 ? 32768 : 0)): round to nearest even, negatives saturate to 0;
     * clamp at 65535 (the reference would overflow into the other half, gpu_kernel.cu:410-412). */
    float v = x + (x != 0.0f ? 32768.0f : 0.0f);
    if (!(v > 0.0f)) return 0u; /* also NaN -> 0, as the hardware conversion does */
    if (v >= 65535.0f) return 65535u;
    return (uint32_t)nearbyintf(v);
}

uint32_t orc_encode_u16_pair(float re, float im)
{
    return (enc16(im) << 16) | enc16(re);
}

/* ------------------------------------------------------------------------------------------------
 * Seeded synthetic pulsar (SplitMix64 + Box-Muller; dispersion with the direct chirp).
 * --------------------------------------------------------------------------------------------- */
static uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t *s)
{
    return ((double)(sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static void gauss2(uint64_t *s, double *a, double *b)
{
    double u1 = u01(s), u2 = u01(s);
    double r = sqrt(-2.0 * log(u1));
    *a = r * cos(2.0 * M_PI * u2);
    *b = r * sin(2.0 * M_PI * u2);
}

int orc_pulsar(long n, double fs, double f0, double dm, double period, double duty, double amp,
               double sigma, uint64_t seed, double *out)
{
    if (n < 2 || (n & (n - 1)) != 0) return -1;
    uint64_t s = seed;
    const double sig_p = duty / (2.0 * sqrt(2.0 * log(2.0))); /* FWHM -> std, in phase units */
    const double cs = sqrt(0.5);
    for (long i = 0; i < n; i++) {
        double t = (double)i / fs;
        double ph = fmod(t / period, 1.0) - 0.5;
        double a = amp * exp(-0.5 * (ph / sig_p) * (ph / sig_p));
        double g1, g2;
        gauss2(&s, &g1, &g2);
        out[2 * i] = a * g1 * cs;
        out[2 * i + 1] = a * g2 * cs;
    }
    if (dm != 0.0) {
        if (orc_fft(out, n, -1) != 0) return -1;
        const double kdm_d = 4.15e15, w0 = 2.0 * M_PI * f0;
        for (long k = 0; k < n; k++) {
            double f = (k < n / 2) ? fs / 2.0 + (double)k * fs / (double)n : (double)k * fs / (double)n - fs / 2.0;
            double w = 2.0 * M_PI * f;
            double phase = 4.0 * M_PI * M_PI * kdm_d * dm * w * w / ((w + w0) * w0 * w0);
            double c = cos(phase), sn = sin(phase);
            double re = out[2 * k], im = out[2 * k + 1];
            out[2 * k] = (re * c - im * sn) / (double)n;
            out[2 * k + 1] = (re * sn + im * c) / (double)n;
        }
        if (orc_fft(out, n, +1) != 0) return -1;
    }
    for (long i = 0; i < n; i++) {
        double g1, g2;
        gauss2(&s, &g1, &g2);
        out[2 * i] += sigma * g1 * cs;
        out[2 * i + 1] += sigma * g2 * cs;
    }
    return 0;
}
