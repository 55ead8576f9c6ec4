// Uneven DM grids: every DM keeps its own MS-OSM length. DMs with the same length form a class that
// runs as one device plan; all classes share the plan stream and one input stream, so one push
// dedisperses the same samples for every DM. There is no reference counterpart (the reference uses
// one common M for all DMs, msosm_prepare.cpp:67); outputs and folds follow the single-plan layouts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "msosm_amd.h"
#include "plan_internal.h"

using namespace msosm;

#define HIPCHK_MULTI(x)                                                                                                \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
    } while (0)

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
    // one dedispersion launch over all classes (MSOSM_LAUNCH_PERSISTENT), when every class fits it
    int launch_mode = MSOSM_LAUNCH_PERSISTENT;
    unsigned int *counter = nullptr;
    int max_fused_bins = 0;
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    double prof_ms = 0;
    long long prof_n = 0;
    // hipGraph form of the one-launch push: built on the first push, then re-parameterised per push
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> n_fwd, n_hist, n_bins;
    hipGraphNode_t n_fill = nullptr, n_inv = nullptr;
    std::vector<PushArgs> gargs;
    MultiInvArgs gma{};
    unsigned int *fill_ptr = nullptr;
    long long fill_words = 1;
    uint32_t fill_value = 0;
};

static void drop_graph(msosm_multi *m)
{
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
    m->n_fwd.clear();
    m->n_hist.clear();
    m->n_bins.clear();
    m->n_fill = m->n_inv = nullptr;
}

// every class can join the one-launch form: lengths 2^14..2^21, rings below 4 GiB, <= MAX_CLASSES
static bool fused_eligible(const msosm_multi *m)
{
    if (m->classes.size() > (size_t)MAX_CLASSES) return false;
    if (m->nbins > (unsigned long)m->max_fused_bins) return false;
    for (const auto &c : m->classes)
        if (!multi_length_supported(plan_log2L(c.plan)) || !plan_ring32(c.plan)) return false;
    return true;
}

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
    drop_graph(m);
    if (m->counter) (void)hipFree(m->counter);
    for (auto &e : m->pending) m->pool.push_back(e);
    for (auto &e : m->pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
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
    if (e == hipSuccess) e = hipMalloc(&m->counter, sizeof(unsigned int));
    int max_lds = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return fail(MSOSM_ERR_ALLOC, std::string("multi_create: ") + hipGetErrorString(e));
    const int tiles = multi_tiles_lds_bytes();
    m->max_fused_bins = std::max(0, std::min(2048, (max_lds - tiles - INV_EXTRA_LDS) / 12));
    e = configure_multi_lds(tiles + m->max_fused_bins * 12 + INV_EXTRA_LDS);
    if (e != hipSuccess) return fail(MSOSM_ERR_HIP, std::string("multi_create: LDS attribute: ") + hipGetErrorString(e));
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
    if (mode != MSOSM_LAUNCH_PERSISTENT && mode != MSOSM_LAUNCH_PER_DM && mode != MSOSM_LAUNCH_PER_CLASS)
        return set_error(MSOSM_ERR_ARG, "multi_set_launch_mode: unknown mode");
    for (auto &c : m->classes) {
        const int rc = msosm_plan_set_launch_mode(c.plan, mode == MSOSM_LAUNCH_PER_DM ? MSOSM_LAUNCH_PER_DM
                                                                                     : MSOSM_LAUNCH_PERSISTENT);
        if (rc != MSOSM_OK) return rc;
    }
    m->launch_mode = mode;
    drop_graph(m);
    return MSOSM_OK;
}

extern "C" int msosm_multi_set_graph(msosm_multi *m, int enable)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_set_graph: null object");
    HIPCHK_MULTI(hipStreamSynchronize(m->stream));
    drop_graph(m);
    m->use_graph = enable != 0;
    return MSOSM_OK;
}

static void multi_inv_args(msosm_multi *m, MultiInvArgs &ma, const std::vector<InvArgs> &inv)
{
    ma = MultiInvArgs{};
    long long acc = 0;
    const int n = (int)m->classes.size();
    for (int i = 0; i < n; ++i) {
        ma.cls[i] = inv[i];
        ma.log2L[i] = plan_log2L(m->classes[i].plan);
        ma.ustart[i] = acc;
        acc += plan_num_units(m->classes[i].plan);
    }
    ma.ustart[n] = acc;
    ma.ncls = n;
    ma.nbins = ma.cls[0].fold_sum ? ma.cls[0].nbins : 0;
    // class 0 has the longest transforms; its workgroup scratch slots hold any class's
    ma.scratch = ma.cls[0].scratch;
    ma.scr_stride = 1LL << ma.log2L[0];
    ma.counter = m->counter;
}

static hipKernelNodeParams node_params(const NodeDesc &d, void **args)
{
    hipKernelNodeParams p{};
    p.func = const_cast<void *>(d.func);
    p.gridDim = d.grid;
    p.blockDim = d.block;
    p.sharedMemBytes = d.lds;
    p.kernelParams = args;
    p.extra = nullptr;
    return p;
}

// The one-launch push as a hipGraph: [counter fill] + per class [forward -> history, bins] ->
// multi-length dedispersion. The graph is built once; later pushes only update the node arguments
// (ring positions, input address, fold clock) and launch it.
static int graph_push(msosm_multi *m, const void *src)
{
    const int n = (int)m->classes.size();
    m->gargs.resize(n);
    std::vector<InvArgs> inv(n);
    for (int i = 0; i < n; ++i) {
        plan_push_args(m->classes[i].plan, src, &m->gargs[i], false);
        inv[i] = m->gargs[i].inv;
    }
    multi_inv_args(m, m->gma, inv);
    m->fill_ptr = m->counter;
    m->fill_words = 1;
    m->fill_value = 0;
    const int grid = plan_workgroups(m->classes[0].plan);
    // argument pointer lists (hipGraph*KernelNode* copy the values they point to)
    std::vector<std::array<void *, 6>> pf(n), ph(n), pb(n);
    for (int i = 0; i < n; ++i) {
        PushArgs &a = m->gargs[i];
        pf[i] = {&a.fwd, nullptr, nullptr, nullptr, nullptr, nullptr};
        ph[i] = {&a.hist.hist, &a.hist.in, &a.hist.fmt, &a.hist.first, &a.hist.M, nullptr};
        pb[i] = {&a.bins.bins, &a.bins.n, &a.bins.nbins, &a.bins.t_start, &a.bins.dt, &a.bins.period};
    }
    void *pfill[3] = {&m->fill_ptr, &m->fill_words, &m->fill_value};
    void *pinv[1] = {&m->gma};
    const NodeDesc dfill = fill_node(4), dinv = inv_multi_node(m->gma, grid);
    if (!m->exec) {
        HIPCHK_MULTI(hipGraphCreate(&m->graph, 0));
        hipKernelNodeParams kp = node_params(dfill, pfill);
        HIPCHK_MULTI(hipGraphAddKernelNode(&m->n_fill, m->graph, nullptr, 0, &kp));
        std::vector<hipGraphNode_t> deps{m->n_fill};
        m->n_fwd.assign(n, nullptr);
        m->n_hist.assign(n, nullptr);
        m->n_bins.assign(n, nullptr);
        for (int i = 0; i < n; ++i) {
            const PushArgs &a = m->gargs[i];
            kp = node_params(fwd_node(a.log2L, a.fwd_grid), pf[i].data());
            HIPCHK_MULTI(hipGraphAddKernelNode(&m->n_fwd[i], m->graph, nullptr, 0, &kp));
            kp = node_params(hist_node(a.hist.M), ph[i].data());
            HIPCHK_MULTI(hipGraphAddKernelNode(&m->n_hist[i], m->graph, &m->n_fwd[i], 1, &kp));
            deps.push_back(m->n_fwd[i]);
            if (a.has_bins) {
                kp = node_params(bins_node(a.bins.n), pb[i].data());
                HIPCHK_MULTI(hipGraphAddKernelNode(&m->n_bins[i], m->graph, nullptr, 0, &kp));
                deps.push_back(m->n_bins[i]);
            }
        }
        kp = node_params(dinv, pinv);
        HIPCHK_MULTI(hipGraphAddKernelNode(&m->n_inv, m->graph, deps.data(), deps.size(), &kp));
        HIPCHK_MULTI(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    } else {
        for (int i = 0; i < n; ++i) {
            const PushArgs &a = m->gargs[i];
            hipKernelNodeParams kp = node_params(fwd_node(a.log2L, a.fwd_grid), pf[i].data());
            HIPCHK_MULTI(hipGraphExecKernelNodeSetParams(m->exec, m->n_fwd[i], &kp));
            kp = node_params(hist_node(a.hist.M), ph[i].data());
            HIPCHK_MULTI(hipGraphExecKernelNodeSetParams(m->exec, m->n_hist[i], &kp));
            if (a.has_bins) {
                kp = node_params(bins_node(a.bins.n), pb[i].data());
                HIPCHK_MULTI(hipGraphExecKernelNodeSetParams(m->exec, m->n_bins[i], &kp));
            }
        }
        hipKernelNodeParams kp = node_params(dinv, pinv);
        HIPCHK_MULTI(hipGraphExecKernelNodeSetParams(m->exec, m->n_inv, &kp));
    }
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (m->profile) {
        if (m->pool.empty()) {
            HIPCHK_MULTI(hipEventCreate(&ev.first));
            HIPCHK_MULTI(hipEventCreate(&ev.second));
        } else {
            ev = m->pool.back();
            m->pool.pop_back();
        }
        HIPCHK_MULTI(hipEventRecord(ev.first, m->stream));
    }
    HIPCHK_MULTI(hipGraphLaunch(m->exec, m->stream));
    if (m->profile) {
        HIPCHK_MULTI(hipEventRecord(ev.second, m->stream));
        m->pending.push_back(ev);
    }
    for (auto &c : m->classes) plan_push_commit(c.plan);
    return MSOSM_OK;
}

// One dedispersion launch over every class: the forward parts of all class pushes, then the
// multi-length persistent kernel, then each class's ring/fold advance.
static int fused_push(msosm_multi *m, const void *src)
{
    const int n = (int)m->classes.size();
    std::vector<InvArgs> inv(n);
    for (int i = 0; i < n; ++i) {
        const int rc = plan_push_prepare(m->classes[i].plan, src, &inv[i], false);
        if (rc != MSOSM_OK) return rc;
    }
    MultiInvArgs ma;
    multi_inv_args(m, ma, inv);
    hipError_t e = launch_fill(m->counter, sizeof(unsigned int), 0u, m->stream);
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (e == hipSuccess && m->profile) {
        if (m->pool.empty()) {
            e = hipEventCreate(&ev.first);
            if (e == hipSuccess) e = hipEventCreate(&ev.second);
        } else {
            ev = m->pool.back();
            m->pool.pop_back();
        }
        if (e == hipSuccess) e = hipEventRecord(ev.first, m->stream);
    }
    if (e == hipSuccess) e = launch_inv_multi(ma, plan_workgroups(m->classes[0].plan), m->stream);
    if (e == hipSuccess && m->profile) {
        e = hipEventRecord(ev.second, m->stream);
        m->pending.push_back(ev);
    }
    if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("multi_push: ") + hipGetErrorString(e));
    for (auto &c : m->classes) plan_push_commit(c.plan);
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
    if (m->launch_mode == MSOSM_LAUNCH_PERSISTENT && fused_eligible(m)) {
        for (auto &c : m->classes)
            if (plan_check_ready(c.plan) != MSOSM_OK) return MSOSM_ERR_STATE;
        const int rc = m->use_graph ? graph_push(m, src) : fused_push(m, src);
        if (rc != MSOSM_OK) return rc;
    } else {
        for (auto &c : m->classes) {
            const int rc = msosm_plan_push(c.plan, src, MSOSM_INPUT_DEVICE);
            if (rc != MSOSM_OK) return rc;
        }
    }
    if (mode != MSOSM_INPUT_DEVICE) {
        hipError_t e = hipEventRecord(m->stage_ev[slot], m->stream);
        if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("multi_push: ") + hipGetErrorString(e));
    }
    m->pushes += 1;
    return MSOSM_OK;
}

// every class plan and fold back to the stream start (zero history, empty folds)
static int multi_reset(msosm_multi *m)
{
    for (auto &c : m->classes) {
        int rc = msosm_plan_reset(c.plan);
        if (rc == MSOSM_OK && c.fold) rc = msosm_fold_reset(c.fold);
        if (rc != MSOSM_OK) return rc;
    }
    return MSOSM_OK;
}

// The autotuner of an uneven grid: each length class's plan tunes its own tile schedule (unit order,
// workgroups, launch form: msosm_plan_autotune on the same device input), then the push is timed as one
// launch over all classes and as one tuned launch per class, and the faster is kept. The tuning pushes
// consume the stream: every class plan and fold is reset afterwards.
extern "C" int msosm_multi_autotune(msosm_multi *m, const void *input, int pushes, int *best_mode, double *best_ms)
{
    if (!m || !input) return set_error(MSOSM_ERR_ARG, "multi_autotune: null argument");
    if (pushes < 2) return set_error(MSOSM_ERR_ARG, "multi_autotune: need at least 2 pushes of input");
    if (m->launch_mode == MSOSM_LAUNCH_PER_DM)
        return set_error(MSOSM_ERR_STATE, "multi_autotune: tunes the persistent and per-class forms only");
    HIPCHK_MULTI(hipSetDevice(m->device));
    const bool graph = m->use_graph;
    m->use_graph = false;
    drop_graph(m);
    for (auto &c : m->classes) {
        int w = 0, o = 0;
        double t = 0;
        const int rc = msosm_plan_autotune(c.plan, input, pushes, &w, &o, &t);
        if (rc != MSOSM_OK) return rc;
    }
    const size_t step = (size_t)m->push * m->in_bytes;
    double best = 1e300;
    int mode = MSOSM_LAUNCH_PER_CLASS;
    hipEvent_t a = nullptr, b = nullptr;
    HIPCHK_MULTI(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    int rc = e == hipSuccess ? MSOSM_OK : set_error(MSOSM_ERR_HIP, std::string("multi_autotune: ") + hipGetErrorString(e));
    for (int cand : {MSOSM_LAUNCH_PERSISTENT, MSOSM_LAUNCH_PER_CLASS}) {
        if (rc != MSOSM_OK) break;
        if (cand == MSOSM_LAUNCH_PERSISTENT && !fused_eligible(m)) continue;
        m->launch_mode = cand;
        rc = multi_reset(m);
        // the first push of each candidate is not timed
        if (rc == MSOSM_OK) rc = msosm_multi_push(m, input, MSOSM_INPUT_DEVICE);
        if (rc == MSOSM_OK && hipEventRecord(a, m->stream) != hipSuccess) rc = set_error(MSOSM_ERR_HIP, "multi_autotune: event");
        for (int k = 1; k < pushes && rc == MSOSM_OK; ++k)
            rc = msosm_multi_push(m, static_cast<const char *>(input) + k * step, MSOSM_INPUT_DEVICE);
        float ms = 0;
        if (rc == MSOSM_OK) {
            e = hipEventRecord(b, m->stream);
            if (e == hipSuccess) e = hipEventSynchronize(b);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
            if (e != hipSuccess) rc = set_error(MSOSM_ERR_HIP, std::string("multi_autotune: ") + hipGetErrorString(e));
        }
        const double per = ms / (pushes - 1);
        if (rc == MSOSM_OK && per < best) {
            best = per;
            mode = cand;
        }
    }
    (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc != MSOSM_OK) return rc;
    m->launch_mode = mode;
    m->use_graph = graph;
    rc = multi_reset(m);
    if (rc != MSOSM_OK) return rc;
    if (best_mode) *best_mode = mode;
    if (best_ms) *best_ms = best;
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
    HIPCHK_MULTI(hipStreamSynchronize(m->stream));
    drop_graph(m);
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

static int collect_fused(msosm_multi *m)
{
    for (auto &ev : m->pending) {
        float ms = 0;
        hipError_t e = hipEventSynchronize(ev.second);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev.first, ev.second);
        if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("multi_profile: ") + hipGetErrorString(e));
        m->prof_ms += ms;
        m->prof_n += 1;
        m->pool.push_back(ev);
    }
    m->pending.clear();
    return MSOSM_OK;
}

extern "C" int msosm_multi_profile(msosm_multi *m, int enable)
{
    if (!m) return set_error(MSOSM_ERR_ARG, "multi_profile: null object");
    if (collect_fused(m) != MSOSM_OK) return MSOSM_ERR_HIP;
    m->profile = enable != 0;
    m->prof_ms = 0;
    m->prof_n = 0;
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
    if (kernel == 1) {
        if (collect_fused(m) != MSOSM_OK) return MSOSM_ERR_HIP;
        ms += m->prof_ms;
        n += m->prof_n;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
    return MSOSM_OK;
}
