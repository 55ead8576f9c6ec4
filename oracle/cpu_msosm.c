/*
 * FP32 CPU port of MS-OSM + fold -- TEST INFRASTRUCTURE ONLY (the timed CPU baseline).
 *
 * The same algorithm as orc_msosm + orc_fold (msosm_oracle.c), in FP32 arithmetic with the
 * transforms done through the FFTW3 single-precision API, as the reference's own CPU overlap-save
 * checks do (benchmark/check_osm.cpp:73-154, benchmark/check_msosm.cpp:182-219: fftwf plans over
 * blocks of 2M samples). No FFTW library is installed here; the FFTW3 interface of Intel MKL
 * (libmkl_rt, fftwf_plan_dft_1d / fftwf_execute_dft) is loaded at run time and used in its
 * sequential threading layer, one plan per OpenMP thread. Without MKL the port falls back to its own
 * iterative radix-2 FP32 FFT. The backend in use is reported to the caller.
 *
 * Steps (restating the reference path, msosm_gpu_batch.cpp:153-201 and gpu_kernel.cu:105-210,
 * 346-376, 462-496):
 *   1. X_g = DFT_2M(x[(g-1)M, (g+1)M)) for every block g (x before sample 0 is zero);
 *   2. per DM d and block g: Y[k] = X_{g - delay_d(k)}[k] * H_d[k] (X before block 0 is zero),
 *      y = IDFT_2M(Y), out[gM + m] = y[M + m] / (2M);
 *   3. fold |out|^2 into FP64 sums and counts at the reference's bin of each sample (zeros skipped).
 * Parallel over DM trials (each DM's fold is owned by one thread).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef float cfloat[2];
typedef void *fplan;

static struct {
    int tried, mkl; /* MKL's FFTW3 interface loaded */
    fplan (*plan_dft_1d)(int, cfloat *, cfloat *, int, unsigned);
    void (*execute_dft)(const fplan, cfloat *, cfloat *);
    void (*destroy_plan)(fplan);
    char name[160];
} B;

#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)
#define FFTW_ESTIMATE (1U << 6)
#define MKL_THREADING_SEQUENTIAL 1

static const char *OWN_NAME = "own radix-2 FP32 FFT";

static void backend_init(void)
{
    if (B.tried) return;
    B.tried = 1;
    snprintf(B.name, sizeof B.name, "%s", OWN_NAME);
    void *h = dlopen("libmkl_rt.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("/opt/conda/lib/libmkl_rt.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
    void (*version)(char *, int) = (void (*)(char *, int))dlsym(h, "MKL_Get_Version_String");
    B.plan_dft_1d = (fplan(*)(int, cfloat *, cfloat *, int, unsigned))dlsym(h, "fftwf_plan_dft_1d");
    B.execute_dft = (void (*)(const fplan, cfloat *, cfloat *))dlsym(h, "fftwf_execute_dft");
    B.destroy_plan = (void (*)(fplan))dlsym(h, "fftwf_destroy_plan");
    if (!set_layer || !B.plan_dft_1d || !B.execute_dft || !B.destroy_plan) return;
    set_layer(MKL_THREADING_SEQUENTIAL); /* before any other MKL call: sequential transforms */
    char v[128] = "";
    if (version) version(v, (int)sizeof v);
    B.mkl = 1;
    snprintf(B.name, sizeof B.name, "FFTW3 API (fftwf_plan_dft_1d) of %s", v[0] ? v : "Intel MKL");
}

/* ---- own FFT: iterative radix-2 (bit reversal + butterflies), FP32, twiddles from FP64 ---- */
typedef struct {
    int n, lg;
    float *tw; /* e^{-2 pi i m / n}, m < n/2 */
} own_fft;

static int own_init(own_fft *f, int n)
{
    f->n = n;
    f->lg = 0;
    while ((1 << f->lg) < n) ++f->lg;
    f->tw = (float *)malloc(sizeof(float) * (size_t)n);
    if (!f->tw) return -1;
    for (int m = 0; m < n / 2; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)n;
        f->tw[2 * m] = (float)cos(a);
        f->tw[2 * m + 1] = (float)sin(a);
    }
    return 0;
}

static void own_run(const own_fft *f, float *x, int sign)
{
    const int n = f->n;
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            float t0 = x[2 * i], t1 = x[2 * i + 1];
            x[2 * i] = x[2 * j];
            x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = t0;
            x[2 * j + 1] = t1;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; ++k) {
                const float wr = f->tw[2 * k * step], wi = -sign * f->tw[2 * k * step + 1]; /* table is e^{-i} */
                float *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                const float br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br;
                b[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
    }
}

/* per-thread transform of length L on buffer buf, sign -1 (forward) or +1 (inverse, unnormalised) */
typedef struct {
    int mkl;
    fplan fwd, inv;
    own_fft own;
    float *buf;
} worker;

static int worker_init(worker *w, int L, int mkl)
{
    memset(w, 0, sizeof *w);
    w->mkl = mkl;
    if (posix_memalign((void **)&w->buf, 64, sizeof(float) * 2 * (size_t)L)) return -1;
    if (mkl) {
        w->fwd = B.plan_dft_1d(L, (cfloat *)w->buf, (cfloat *)w->buf, FFTW_FORWARD, FFTW_ESTIMATE);
        w->inv = B.plan_dft_1d(L, (cfloat *)w->buf, (cfloat *)w->buf, FFTW_BACKWARD, FFTW_ESTIMATE);
        return (w->fwd && w->inv) ? 0 : -1;
    }
    return own_init(&w->own, L);
}

static void worker_run(worker *w, int sign)
{
    if (w->mkl)
        B.execute_dft(sign < 0 ? w->fwd : w->inv, (cfloat *)w->buf, (cfloat *)w->buf);
    else
        own_run(&w->own, w->buf, sign);
}

static void worker_free(worker *w)
{
    if (w->mkl) {
        if (w->fwd) B.destroy_plan(w->fwd);
        if (w->inv) B.destroy_plan(w->inv);
    }
    free(w->own.tw);
    free(w->buf);
}

/* reference fold bin (gpu_kernel.cu:477-481), the same expression as orc_fold_bin */
static long fold_bin(long i, unsigned long nbins, double t_start, double dt, double period)
{
    const double t = t_start + (double)i * dt; /* built with -ffp-contract=off: no FMA */
    const double phase = fmod(t / period, 1.0);
    unsigned long b = (unsigned long)floor(phase * (double)nbins);
    if (b >= nbins) b = nbins - 1;
    return (long)b;
}

/* name of the transform backend a call with this force_own value uses */
const char *cpu_msosm_backend(int force_own)
{
    backend_init();
    return (B.mkl && !force_own) ? B.name : OWN_NAME;
}

/* The record is folded as calls of call_size samples (0: one call), as Fold_GPU::fold_data_phase folds
 * each call's series: call c bins its sample i at t_start_c + i*dt with t_start_{c+1} = t_start_c +
 * call_size*dt (fold_gpu.cpp:48-54). */
int cpu_msosm_fold(const float *x, long n, int M, int D, const uint16_t *delays, const float *H, unsigned long nbins,
                   double t_start, double dt, double period, long call_size, double *sums, unsigned long long *cnts,
                   float *out, int nthreads, int force_own)
{
    if (!x || n <= 0 || M <= 0 || (n % M) != 0 || D <= 0 || !delays || !H || nbins == 0 || nbins > 65536 || !sums ||
        !cnts)
        return -1;
    backend_init();
    const int mkl = B.mkl && !force_own;
    const int L = 2 * M;
    const long nblk = n / M;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const long call = call_size > 0 ? call_size : n;
    const long ncalls = (n + call - 1) / call;
    double *ts = (double *)malloc(sizeof(double) * (size_t)ncalls);
    if (!ts) return -1;
    ts[0] = t_start;
    for (long c = 1; c < ncalls; ++c) ts[c] = ts[c - 1] + (double)call * dt;
    float *spec = NULL;
    if (posix_memalign((void **)&spec, 64, sizeof(float) * 2 * (size_t)L * (size_t)nblk)) spec = NULL;
    uint16_t *bins = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n);
    worker *ws = (worker *)calloc((size_t)nthreads, sizeof(worker));
    int rc = (spec && bins && ws) ? 0 : -1;
    for (int t = 0; rc == 0 && t < nthreads; ++t) rc = worker_init(&ws[t], L, mkl); /* planning is serial */
    if (rc == 0) {
        const float scale = 1.0f / (float)L;
#pragma omp parallel num_threads(nthreads)
        {
            worker *w = &ws[omp_get_thread_num()];
#pragma omp for schedule(static)
            for (long i = 0; i < n; ++i) bins[i] = (uint16_t)fold_bin(i % call, nbins, ts[i / call], dt, period);
            /* 1. spectra of every block */
#pragma omp for schedule(static)
            for (long g = 0; g < nblk; ++g) {
                for (long s = 0; s < L; ++s) {
                    const long j = (g - 1) * M + s;
                    w->buf[2 * s] = j < 0 ? 0.f : x[2 * j];
                    w->buf[2 * s + 1] = j < 0 ? 0.f : x[2 * j + 1];
                }
                worker_run(w, -1);
                memcpy(spec + 2 * (size_t)L * g, w->buf, sizeof(float) * 2 * (size_t)L);
            }
            /* 2.-3. per DM: gather * H, inverse transform, discard, fold */
#pragma omp for schedule(dynamic, 1)
            for (int d = 0; d < D; ++d) {
                const uint16_t *dl = delays + (size_t)d * L;
                const float *h = H + 2 * (size_t)d * L;
                double *ds = sums + (size_t)d * nbins;
                unsigned long long *dc = cnts + (size_t)d * nbins;
                for (long g = 0; g < nblk; ++g) {
                    for (int k = 0; k < L; ++k) {
                        const long src = g - (long)dl[k];
                        if (src < 0) {
                            w->buf[2 * k] = w->buf[2 * k + 1] = 0.f;
                            continue;
                        }
                        const float *X = spec + 2 * ((size_t)L * src + k);
                        w->buf[2 * k] = X[0] * h[2 * k] - X[1] * h[2 * k + 1];
                        w->buf[2 * k + 1] = X[0] * h[2 * k + 1] + X[1] * h[2 * k];
                    }
                    worker_run(w, +1);
                    for (int m = 0; m < M; ++m) {
                        const float yr = w->buf[2 * (M + m)] * scale, yi = w->buf[2 * (M + m) + 1] * scale;
                        const long i = g * M + m;
                        if (out) {
                            out[2 * ((size_t)d * n + i)] = yr;
                            out[2 * ((size_t)d * n + i) + 1] = yi;
                        }
                        const float I = yr * yr + yi * yi;
                        if (I == 0.0f) continue;
                        ds[bins[i]] += (double)I;
                        dc[bins[i]] += 1;
                    }
                }
            }
        }
    }
    if (ws)
        for (int t = 0; t < nthreads; ++t) worker_free(&ws[t]);
    free(ws);
    free(bins);
    free(spec);
    free(ts);
    return rc;
}
