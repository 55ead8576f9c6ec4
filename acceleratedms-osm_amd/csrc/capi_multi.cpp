// Uneven DM grids: every DM keeps its own MS-OSM length. DMs with the same length form a class that
// runs as one device plan; all classes share the plan stream and one input stream, so one push
// dedisperses the same samples for every DM. There is no reference counterpart (the reference uses
// one common M for all DMs, msosm_prepare.cpp:67); outputs and folds follow the single-plan layouts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "msosm_amd.h"

using namespace msosm;

namespace {

struct Class {
    int M = 0;
    std::vector<int> dms; // global DM indices, in plan order
    msosm_plan *plan = nullptr;
    msosm_fold *fold = nullptr;
};

} // namespace

struct msosm_multi {
    int D = 0, in_fmt = 0, out_flags = 0, device = 0;
    long long push = 0; // samples per push
    size_t in_bytes = 0;
    std::vector<Class> classes;
    std::vector<int> cls_of, local_of; // per global DM
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    long long pushes = 0;
    hipStream_t stream = nullptr;
    unsigned long nbins = 0;
};

static void multi_free(msosm_multi *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (auto &c : m->classes) {
        if (c.fold) msosm_fold_destroy(c.fold);
    }
    // plans: destroy the stream owner (the first) last
    for (size_t i = m->classes.size(); i-- > 0;)
        if (m->classes[i].plan) msosm_plan_destroy(m->classes[i].plan);
    for (int i = 0; i < 2; ++i) {
        if (m->stage[i]) (void)hipFree(m->stage[i]);
        if (m->stage_ev[i]) (void)hipEventDestroy(m->stage_ev[i]);
    }
    delete m;
}

extern "C" int msosm_design_uneven_order(int nd, int *M)
{
    // the reference's per-DM minimum order (msosm_prepare.cpp:59-61), clamped to FFT lengths 2^14..2^21
    if (!M || nd < 0) return set_error(MSOSM_ERR_ARG, "design_uneven_order: nd must be >= 0");
    int m = nd == 0 ? 0 : design_min_order(nd);
    *M = std::min(1 << 20, std::max(1 << 13, m));
    return MSOSM_OK;
}

extern "C" int msosm_multi_create(const msosm_multi_desc *d, msosm_multi **out)
{
    if (!d || !out || d->num_dms <= 0 || !d->M || !d->delaycount || d->push_blocks <= 0)
        return set_error(MSOSM_ERR_ARG, "multi_create: bad descriptor");
    *out = nullptr;
    msosm_multi *m = new msosm_multi();
    auto fail = [&](int code, const std::string &msg) {
        multi_free(m);
        return set_error(code, msg);
    };
    m->D = d->num_dms;
    m->in_fmt = d->in_format;
    m->out_flags = d->out_flags;
    m->in_bytes = d->in_format == MSOSM_IN_U16 ? 4 : d->in_format == MSOSM_IN_I8 ? 2 : 8;
    int dev = d->device;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) dev = 0;
    m->device = dev;
    int Mmax = 0;
    std::map<int, std::vector<int>> by_M;
    std::map<int, int> dc_of;
    for (int i = 0; i < m->D; ++i) {
        const int Mi = d->M[i];
        if (Mi <= 0 || (Mi & (Mi - 1))) return fail(MSOSM_ERR_ARG, "multi_create: every M must be a power of two");
        by_M[Mi].push_back(i);
        dc_of[Mi] = std::max(dc_of[Mi], d->delaycount[i]);
        Mmax = std::max(Mmax, Mi);
    }
    m->push = (long long)d->push_blocks * Mmax;
    m->cls_of.assign(m->D, -1);
    m->local_of.assign(m->D, -1);
    // largest lengths first: the first plan owns the stream, the others run on it
    for (auto it = by_M.rbegin(); it != by_M.rend(); ++it) {
        Class c;
        c.M = it->first;
        c.dms = it->second;
        msosm_plan_desc pd{};
        pd.M = c.M;
        pd.num_dms = (int)c.dms.size();
        pd.batch = (int)(m->push / c.M);
        pd.delaycount = dc_of[c.M];
        pd.in_format = d->in_format;
        pd.out_flags = d->out_flags;
        pd.device = dev;
        pd.stream = m->classes.empty() ? d->stream : (void *)m->stream;
        int rc = msosm_plan_create(&pd, &c.plan);
        if (rc != MSOSM_OK) {
            std::string msg = "multi_create: class M=" + std::to_string(c.M) + ": " + msosm_last_error();
            return fail(rc, msg);
        }
        if (m->classes.empty()) {
            void *st = nullptr;
            msosm_plan_get_stream(c.plan, &st);
            m->stream = static_cast<hipStream_t>(st);
        }
        const int ci = (int)m->classes.size();
        for (size_t j = 0; j < c.dms.size(); ++j) {
            m->cls_of[c.dms[j]] = ci;
            m->local_of[c.dms[j]] = (int)j;
        }
        m->classes.push_back(c);
    }
    hipError_t e = hipSetDevice(dev);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipMalloc(&m->stage[i], (size_t)m->push * m->in_bytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&m->stage_ev[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) return fail(MSOSM_ERR_ALLOC, std::string("multi_create: ") + hipGetErrorString(e));
    *out = m;
    return MSOSM_OK;
}

extern "C" int msosm_multi_destroy(msosm_multi *m)
{
    multi_free(m);
    return MSOSM_OK;
}

static int check_dm(msosm_multi *m, int dm)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi: null object");
    if (dm < 0 || dm >= m->D) return set_error(MSOSM_ERR_ARG, "multi: dm index out of range");
    return MSOSM_OK;
}

extern "C" int msosm_multi_get_info(msosm_multi *m, int dm, int *M, int *num_classes, long long *push_samples)
{
    if (check_dm(m, dm) != MSOSM_OK) return MSOSM_ERR_ARG;
    if (M) *M = m->classes[m->cls_of[dm]].M;
    if (num_classes) *num_classes = (int)m->classes.size();
    if (push_samples) *push_samples = m->push;
    return MSOSM_OK;
}

extern "C" int msosm_multi_set_dm(msosm_multi *m, int dm, const int *delay_points, const float *H)
{
    if (check_dm(m, dm) != MSOSM_OK) return MSOSM_ERR_ARG;
    return msosm_plan_set_dm(m->classes[m->cls_of[dm]].plan, m->local_of[dm], delay_points, H);
}

extern "C" int msosm_multi_set_launch_mode(msosm_multi *m, int mode)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_set_launch_mode: null object");
    for (auto &c : m->classes) {
        const int rc = msosm_plan_set_launch_mode(c.plan, mode);
        if (rc != MSOSM_OK) return rc;
    }
    return MSOSM_OK;
}

extern "C" int msosm_multi_push(msosm_multi *m, const void *input, int mode)
{
    if (!m || !input) return set_error(MSOSM_ERR_ARG, "multi_push: null argument");
    const void *src = input;
    const int slot = (int)(m->pushes & 1);
    if (mode != MSOSM_INPUT_DEVICE) {
        hipError_t e = hipSetDevice(m->device);
        if (e == hipSuccess) e = hipEventSynchronize(m->stage_ev[slot]);
        if (e == hipSuccess) {
            const size_t bytes = (size_t)m->push * m->in_bytes;
            e = mode == MSOSM_INPUT_HOST_SYNC
                    ? hipMemcpy(m->stage[slot], input, bytes, hipMemcpyHostToDevice)
                    : hipMemcpyAsync(m->stage[slot], input, bytes, hipMemcpyHostToDevice, m->stream);
        }
        if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("multi_push: ") + hipGetErrorString(e));
        src = m->stage[slot];
    }
    for (auto &c : m->classes) {
        const int rc = msosm_plan_push(c.plan, src, MSOSM_INPUT_DEVICE);
        if (rc != MSOSM_OK) return rc;
    }
    if (mode != MSOSM_INPUT_DEVICE) {
        hipError_t e = hipEventRecord(m->stage_ev[slot], m->stream);
        if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("multi_push: ") + hipGetErrorString(e));
    }
    m->pushes += 1;
    return MSOSM_OK;
}

extern "C" int msosm_multi_output_c64(msosm_multi *m, int dm, void **dev_ptr)
{
    if (check_dm(m, dm) != MSOSM_OK || !dev_ptr) return set_error(MSOSM_ERR_ARG, "multi_output_c64: bad arguments");
    void *p = nullptr;
    const int rc = msosm_plan_output_c64(m->classes[m->cls_of[dm]].plan, &p);
    if (rc != MSOSM_OK) return rc;
    *dev_ptr = static_cast<char *>(p) + (size_t)m->local_of[dm] * m->push * 8;
    return MSOSM_OK;
}

extern "C" int msosm_multi_get_output(msosm_multi *m, int dm, void *host)
{
    void *p = nullptr;
    int rc = msosm_multi_output_c64(m, dm, &p);
    if (rc != MSOSM_OK) return rc;
    hipError_t e = hipMemcpyAsync(host, p, (size_t)m->push * 8, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("multi_get_output: ") + hipGetErrorString(e));
    return MSOSM_OK;
}

extern "C" int msosm_multi_fold(msosm_multi *m, double period, double fs, unsigned long nbins)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_fold: null object");
    for (auto &c : m->classes) {
        if (c.fold) {
            msosm_plan_attach_fold(c.plan, nullptr);
            msosm_fold_destroy(c.fold);
            c.fold = nullptr;
        }
        int rc = msosm_fold_create(period, fs, (unsigned long)m->push, nbins, (int)c.dms.size(), m->device, &c.fold);
        if (rc == MSOSM_OK) rc = msosm_plan_attach_fold(c.plan, c.fold);
        if (rc != MSOSM_OK) return rc;
        msosm_fold_get_nbins(c.fold, &m->nbins);
    }
    return MSOSM_OK;
}

extern "C" int msosm_multi_fold_read(msosm_multi *m, float *profile, double *sums, unsigned long long *counts)
{
    if (!m || m->nbins == 0) return set_error(MSOSM_ERR_STATE, "multi_fold_read: no fold attached");
    const size_t nb = m->nbins;
    for (auto &c : m->classes) {
        const size_t n = nb * c.dms.size();
        std::vector<float> p(n);
        std::vector<double> s(n);
        std::vector<unsigned long long> k(n);
        const int rc = msosm_fold_read(c.fold, p.data(), s.data(), k.data());
        if (rc != MSOSM_OK) return rc;
        for (size_t j = 0; j < c.dms.size(); ++j) {
            const size_t g = (size_t)c.dms[j] * nb;
            if (profile) std::copy(p.begin() + j * nb, p.begin() + (j + 1) * nb, profile + g);
            if (sums) std::copy(s.begin() + j * nb, s.begin() + (j + 1) * nb, sums + g);
            if (counts) std::copy(k.begin() + j * nb, k.begin() + (j + 1) * nb, counts + g);
        }
    }
    return MSOSM_OK;
}

extern "C" int msosm_multi_synchronize(msosm_multi *m)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_synchronize: null object");
    return msosm_plan_synchronize(m->classes.front().plan);
}

extern "C" int msosm_multi_profile(msosm_multi *m, int enable)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_profile: null object");
    for (auto &c : m->classes) {
        const int rc = msosm_plan_profile(c.plan, enable);
        if (rc != MSOSM_OK) return rc;
    }
    return MSOSM_OK;
}

extern "C" int msosm_multi_profile_read(msosm_multi *m, int kernel, double *total_ms, long long *launches)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_profile_read: null object");
    double ms = 0;
    long long n = 0;
    for (auto &c : m->classes) {
        double a = 0;
        long long b = 0;
        const int rc = msosm_plan_profile_read(c.plan, kernel, &a, &b);
        if (rc != MSOSM_OK) return rc;
        ms += a;
        n += b;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
    return MSOSM_OK;
}
