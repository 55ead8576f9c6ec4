// Host output pipeline: a ring of slots, each holding one push of per-series output blocks, filled by
// device-to-host copies and written to one file per series by io_uring (raw system calls) or pwrite.
// Restates the slot protocol of MSOSM_DataSave_Uring (data_save/msosm_data_save_uring.cpp:73-288):
// EMPTY -> (producer fills, records an event) COPYING -> (poll thread sees the event) READY ->
// (the slot's writer) SAVING -> EMPTY, and DONE once every expected slot is through. Slot s of push
// k writes series i's block at byte offset k * block_bytes of file i.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <fcntl.h>
#include <linux/io_uring.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "capi_internal.h"
#include "msosm_amd.h"

using namespace msosm;

namespace {

enum SlotState : int { EMPTY = 0, COPYING = 1, READY = 2, SAVING = 3, DONE = 4 };

void cpu_relax(unsigned &spins)
{
    if (++spins < 256)
        __builtin_ia32_pause();
    else
        std::this_thread::yield();
}

// Minimal io_uring: one submission and one completion ring, mapped from the kernel.
struct Uring {
    int fd = -1;
    unsigned *sq_head = nullptr, *sq_tail = nullptr, *sq_mask = nullptr, *sq_array = nullptr;
    unsigned *cq_head = nullptr, *cq_tail = nullptr, *cq_mask = nullptr;
    unsigned sq_entries = 0;
    io_uring_sqe *sqes = nullptr;
    io_uring_cqe *cqes = nullptr;
    void *sq_map = nullptr, *cq_map = nullptr;
    size_t sq_len = 0, cq_len = 0, sqes_len = 0;
    unsigned pending = 0; // sqes written since the last enter

    int init(unsigned entries)
    {
        io_uring_params p;
        memset(&p, 0, sizeof(p));
        fd = (int)syscall(__NR_io_uring_setup, entries, &p);
        if (fd < 0) return -errno;
        sq_entries = p.sq_entries;
        sq_len = p.sq_off.array + p.sq_entries * sizeof(unsigned);
        cq_len = p.cq_off.cqes + p.cq_entries * sizeof(io_uring_cqe);
        const bool single = (p.features & IORING_FEAT_SINGLE_MMAP) != 0;
        if (single) sq_len = cq_len = std::max(sq_len, cq_len);
        sq_map = mmap(nullptr, sq_len, PROT_READ | PROT_WRITE, MAP_SHARED | MAP_POPULATE, fd, IORING_OFF_SQ_RING);
        if (sq_map == MAP_FAILED) return -errno;
        cq_map = single ? sq_map
                        : mmap(nullptr, cq_len, PROT_READ | PROT_WRITE, MAP_SHARED | MAP_POPULATE, fd,
                               IORING_OFF_CQ_RING);
        if (cq_map == MAP_FAILED) return -errno;
        sqes_len = p.sq_entries * sizeof(io_uring_sqe);
        void *s = mmap(nullptr, sqes_len, PROT_READ | PROT_WRITE, MAP_SHARED | MAP_POPULATE, fd, IORING_OFF_SQES);
        if (s == MAP_FAILED) return -errno;
        sqes = static_cast<io_uring_sqe *>(s);
        char *sq = static_cast<char *>(sq_map), *cq = static_cast<char *>(cq_map);
        sq_head = reinterpret_cast<unsigned *>(sq + p.sq_off.head);
        sq_tail = reinterpret_cast<unsigned *>(sq + p.sq_off.tail);
        sq_mask = reinterpret_cast<unsigned *>(sq + p.sq_off.ring_mask);
        sq_array = reinterpret_cast<unsigned *>(sq + p.sq_off.array);
        cq_head = reinterpret_cast<unsigned *>(cq + p.cq_off.head);
        cq_tail = reinterpret_cast<unsigned *>(cq + p.cq_off.tail);
        cq_mask = reinterpret_cast<unsigned *>(cq + p.cq_off.ring_mask);
        cqes = reinterpret_cast<io_uring_cqe *>(cq + p.cq_off.cqes);
        return 0;
    }

    // queue one write; false when the submission ring is full
    bool prep_write(int file, const void *buf, unsigned len, unsigned long long off, unsigned long long tag)
    {
        const unsigned tail = *sq_tail;
        const unsigned head = __atomic_load_n(sq_head, __ATOMIC_ACQUIRE);
        if (tail - head >= sq_entries) return false;
        const unsigned idx = tail & *sq_mask;
        io_uring_sqe *e = &sqes[idx];
        memset(e, 0, sizeof(*e));
        e->opcode = IORING_OP_WRITE;
        e->fd = file;
        e->addr = reinterpret_cast<unsigned long long>(buf);
        e->len = len;
        e->off = off;
        e->user_data = tag;
        sq_array[idx] = idx;
        __atomic_store_n(sq_tail, tail + 1, __ATOMIC_RELEASE);
        ++pending;
        return true;
    }

    // submit what is queued and wait for at least min_complete completions
    int enter(unsigned min_complete)
    {
        for (;;) {
            const int r = (int)syscall(__NR_io_uring_enter, fd, pending, min_complete,
                                       min_complete ? IORING_ENTER_GETEVENTS : 0u, nullptr, 0);
            if (r >= 0) {
                pending -= std::min<unsigned>(pending, (unsigned)r);
                return 0;
            }
            if (errno != EINTR) return -errno;
        }
    }

    bool pop(io_uring_cqe *out)
    {
        const unsigned head = *cq_head;
        const unsigned tail = __atomic_load_n(cq_tail, __ATOMIC_ACQUIRE);
        if (head == tail) return false;
        *out = cqes[head & *cq_mask];
        __atomic_store_n(cq_head, head + 1, __ATOMIC_RELEASE);
        return true;
    }

    void close_ring()
    {
        if (sqes) munmap(sqes, sqes_len);
        if (cq_map && cq_map != MAP_FAILED && cq_map != sq_map) munmap(cq_map, cq_len);
        if (sq_map && sq_map != MAP_FAILED) munmap(sq_map, sq_len);
        if (fd >= 0) ::close(fd);
        fd = -1;
        sqes = nullptr;
        sq_map = cq_map = nullptr;
    }
};

} // namespace

struct msosm_saver {
    std::vector<int> fds;
    std::vector<std::string> paths;
    size_t block = 0, slot_bytes = 0;
    int nslots = 0, flags = 0;
    std::atomic<unsigned long> save_count{0};
    char *slots = nullptr;
    bool pinned = false;
    std::vector<hipEvent_t> events;
    std::vector<char> has_event, bad;
    std::atomic<int> *state = nullptr;
    std::vector<unsigned long long> seq; // push index held by each slot
    int fill_index = 0;
    unsigned long committed = 0;
    std::thread poller;
    std::vector<std::thread> writers;
    std::vector<Uring> rings;
    std::atomic<bool> started{false}; // the loops run (the library's threads, or the caller's: run_*)
    bool own_threads = false;         // msosm_saver_start made them
    std::mutex err_mu;
    std::string err;
    int err_code = MSOSM_OK;
    std::atomic<bool> failed{false};
    bool direct_fallback = false; // O_DIRECT requested but refused for the slot memory
};

static void record_failure(msosm_saver *s, int code, const std::string &msg)
{
    std::lock_guard<std::mutex> g(s->err_mu);
    if (s->err_code == MSOSM_OK) {
        s->err_code = code;
        s->err = msg;
    }
    s->failed.store(true, std::memory_order_release);
}

// Write len bytes of buf at offset off to fd with pwrite, resuming after short writes.
static int write_all_pwrite(int fd, const char *buf, size_t len, unsigned long long off)
{
    while (len > 0) {
        const ssize_t w = pwrite(fd, buf, len, (off_t)off);
        if (w < 0) {
            if (errno == EINTR) continue;
            return -errno;
        }
        if (w == 0) return -EIO;
        buf += w;
        len -= (size_t)w;
        off += (unsigned long long)w;
    }
    return 0;
}

// One slot's blocks (one per file) through io_uring; short writes are resubmitted for the rest. Every
// block is attempted and every completion reaped; the first error is returned.
static int write_slot_uring(msosm_saver *s, Uring &ring, const char *base, unsigned long long off)
{
    const int n = (int)s->fds.size();
    std::vector<size_t> done(n, 0);
    int next = 0, inflight = 0, first_err = 0;
    std::vector<int> resubmit;
    while (next < n || inflight > 0 || !resubmit.empty()) {
        while (!resubmit.empty()) {
            const int i = resubmit.back();
            if (!ring.prep_write(s->fds[i], base + (size_t)i * s->block + done[i], (unsigned)(s->block - done[i]),
                                 off + done[i], (unsigned long long)i))
                break;
            resubmit.pop_back();
            ++inflight;
        }
        while (next < n && ring.prep_write(s->fds[next], base + (size_t)next * s->block, (unsigned)s->block, off,
                                           (unsigned long long)next)) {
            ++next;
            ++inflight;
        }
        const int r = ring.enter(inflight > 0 ? 1u : 0u);
        if (r < 0) return first_err ? first_err : r; // the ring itself failed: nothing more can be reaped
        io_uring_cqe c;
        while (ring.pop(&c)) {
            const int i = (int)c.user_data;
            --inflight;
            if (c.res <= 0) {
                if (!first_err) first_err = c.res < 0 ? c.res : -EIO;
                continue;
            }
            done[i] += (size_t)c.res;
            if (done[i] < s->block) resubmit.push_back(i); // short write: the rest goes in again
        }
    }
    return first_err;
}

static void writer_loop(msosm_saver *s, int k, bool uring)
{
    unsigned spins = 0;
    for (;;) {
        int st = s->state[k].load(std::memory_order_acquire);
        if (st == DONE) return;
        int expected = READY;
        if (st == READY && s->state[k].compare_exchange_strong(expected, SAVING, std::memory_order_acq_rel)) {
            spins = 0;
            // a failed write does not stop the others (each block is attempted); a slot whose copy
            // failed is not written
            if (!s->bad[k]) {
                const char *base = s->slots + (size_t)k * s->slot_bytes;
                const unsigned long long off = s->seq[k] * (unsigned long long)s->block;
                int r = 0;
                if (uring) {
                    r = write_slot_uring(s, s->rings[k], base, off);
                } else {
                    for (size_t i = 0; i < s->fds.size(); ++i) {
                        const int ri = write_all_pwrite(s->fds[i], base + i * s->block, s->block, off);
                        if (ri < 0 && r == 0) r = ri;
                    }
                }
                if (r < 0)
                    record_failure(s, MSOSM_ERR_IO,
                                   std::string("saver: write of push ") + std::to_string(s->seq[k]) + ": " +
                                       strerror(-r));
            }
            s->state[k].store(EMPTY, std::memory_order_release);
        } else {
            cpu_relax(spins);
        }
    }
}

static void poll_loop(msosm_saver *s)
{
    unsigned long polled = 0;
    int k = 0;
    unsigned spins = 0;
    while (polled < s->save_count.load(std::memory_order_acquire)) {
        if (s->state[k].load(std::memory_order_acquire) == COPYING) {
            hipError_t e = s->has_event[k] ? hipEventQuery(s->events[k]) : hipSuccess;
            if (e == hipSuccess) {
                s->state[k].store(READY, std::memory_order_release);
                ++polled;
                k = (k + 1) % s->nslots;
                spins = 0;
                continue;
            }
            if (e != hipErrorNotReady) {
                record_failure(s, MSOSM_ERR_HIP, std::string("saver: hipEventQuery: ") + hipGetErrorString(e));
                s->bad[k] = 1;
                s->state[k].store(READY, std::memory_order_release); // the writer skips it
                ++polled;
                k = (k + 1) % s->nslots;
                continue;
            }
        }
        cpu_relax(spins);
    }
    // every expected slot has been handed over: let each writer finish its last one and stop
    for (int i = 0; i < s->nslots; ++i) {
        while (s->state[i].load(std::memory_order_acquire) != EMPTY) cpu_relax(spins);
        s->state[i].store(DONE, std::memory_order_release);
    }
}

static void saver_free(msosm_saver *s)
{
    if (!s) return;
    for (int fd : s->fds)
        if (fd >= 0) ::close(fd);
    for (auto &r : s->rings) r.close_ring();
    for (size_t i = 0; i < s->events.size(); ++i)
        if (s->events[i]) (void)hipEventDestroy(s->events[i]);
    if (s->slots) {
        if (s->pinned)
            (void)hipHostFree(s->slots);
        else
            free(s->slots);
    }
    delete[] s->state;
    delete s;
}

extern "C" int msosm_saver_file_name(const char *dir, const char *prefix, const float *dms, int num_dms, int i,
                                     char *buf, size_t cap)
{
    // <dir>/<prefix>_DM<value>.dat with the decimal places of the first two DMs
    // (data_save/msosm_data_save_uring.cpp:9-36,56-65)
    if (!dms || num_dms <= 0 || i < 0 || i >= num_dms || !buf || cap == 0)
        return set_error(MSOSM_ERR_ARG, "saver_file_name: bad arguments");
    auto places = [](float dm) {
        dm = std::fabs(dm);
        if (dm <= 0.0f) return 0;
        int ret = 10;
        for (int p = 0; p <= 10; ++p) {
            const float scaled = dm * (float)std::pow(10.0, p);
            if (std::fabs(scaled - std::round(scaled)) < 1e-6f) {
                ret = p;
                break;
            }
        }
        return ret == 0 ? 2 : ret;
    };
    int prec = places(dms[0]);
    if (num_dms > 1) prec = std::max(prec, places(dms[1]));
    const std::string d = (dir && *dir) ? dir : ".";
    const std::string p = (prefix && *prefix) ? std::string(prefix) + "_" : std::string();
    char val[64];
    snprintf(val, sizeof(val), "%.*f", prec, (double)dms[i]);
    const std::string name = d + "/" + p + "DM" + val + ".dat";
    if (name.size() + 1 > cap) return set_error(MSOSM_ERR_ARG, "saver_file_name: buffer too small");
    memcpy(buf, name.c_str(), name.size() + 1);
    return MSOSM_OK;
}

extern "C" int msosm_saver_create(const char *const *paths, int num_files, size_t block_bytes, int slot_count,
                                  unsigned long save_count, int flags, msosm_saver **out)
{
    if (!paths || !out || num_files <= 0 || block_bytes == 0 || slot_count <= 0 || save_count == 0)
        return set_error(MSOSM_ERR_ARG, "saver_create: bad arguments");
    if (block_bytes > 0x7fffffffULL) return set_error(MSOSM_ERR_ARG, "saver_create: block_bytes must be < 2 GiB");
    *out = nullptr;
    msosm_saver *s = new msosm_saver();
    s->block = block_bytes;
    s->slot_bytes = block_bytes * (size_t)num_files;
    s->nslots = slot_count;
    s->save_count.store(save_count);
    s->flags = flags;
    s->state = new std::atomic<int>[slot_count];
    for (int i = 0; i < slot_count; ++i) s->state[i].store(EMPTY);
    s->seq.assign(slot_count, 0);
    s->fds.assign(num_files, -1);
    auto fail = [&](int code, const std::string &m) {
        saver_free(s);
        return set_error(code, m);
    };
    for (int i = 0; i < num_files; ++i) {
        s->paths.emplace_back(paths[i] ? paths[i] : "");
        int oflags = O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC;
        int fd = -1;
        if (flags & MSOSM_SAVE_DIRECT) {
            fd = open(paths[i], oflags | O_DIRECT, 0644);
            // file systems without direct I/O (tmpfs) refuse O_DIRECT with EINVAL: buffered instead
            if (fd < 0 && errno == EINVAL) fd = open(paths[i], oflags, 0644);
        } else {
            fd = open(paths[i], oflags, 0644);
        }
        if (fd < 0) return fail(MSOSM_ERR_IO, std::string("saver_create: open ") + s->paths.back() + ": " + strerror(errno));
        s->fds[i] = fd;
    }
    const size_t total = s->slot_bytes * (size_t)slot_count;
    if (flags & MSOSM_SAVE_PINNED) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&s->slots), total, hipHostMallocDefault);
        if (e != hipSuccess) return fail(MSOSM_ERR_ALLOC, std::string("saver_create: hipHostMalloc: ") + hipGetErrorString(e));
        s->pinned = true;
        s->events.assign(slot_count, nullptr);
        for (int i = 0; i < slot_count; ++i) {
            e = hipEventCreateWithFlags(&s->events[i], hipEventDisableTiming);
            if (e != hipSuccess) return fail(MSOSM_ERR_HIP, std::string("saver_create: ") + hipGetErrorString(e));
        }
    } else {
        // page-aligned so that O_DIRECT writes are aligned too
        if (posix_memalign(reinterpret_cast<void **>(&s->slots), 4096, total) != 0)
            return fail(MSOSM_ERR_ALLOC, "saver_create: host allocation failed");
        s->events.assign(slot_count, nullptr);
    }
    s->has_event.assign(slot_count, 0);
    s->bad.assign(slot_count, 0);
    if (flags & MSOSM_SAVE_DIRECT) {
        // Direct I/O pins the source pages in the kernel; pinned HIP host memory is a device mapping
        // that it cannot pin here (the write fails with EFAULT). Probe one block on the first file
        // and fall back to buffered files when direct writes from the slots are refused.
        const size_t probe = std::min<size_t>(block_bytes, 4096);
        memset(s->slots, 0, probe);
        const ssize_t w = pwrite(s->fds[0], s->slots, probe, 0);
        if (w < 0 && (errno == EFAULT || errno == EINVAL)) {
            s->direct_fallback = true;
            for (int i = 0; i < num_files; ++i) {
                ::close(s->fds[i]);
                s->fds[i] = open(paths[i], O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
                if (s->fds[i] < 0)
                    return fail(MSOSM_ERR_IO, std::string("saver_create: open ") + s->paths[i] + ": " + strerror(errno));
            }
        } else if (w < 0 || ftruncate(s->fds[0], 0) != 0) {
            return fail(MSOSM_ERR_IO, std::string("saver_create: probe write ") + s->paths[0] + ": " + strerror(errno));
        }
    }
    if (flags & MSOSM_SAVE_URING) {
        s->rings.resize(slot_count);
        const unsigned entries = (unsigned)std::min(4096, std::max(8, num_files));
        for (int i = 0; i < slot_count; ++i) {
            const int r = s->rings[i].init(entries);
            if (r < 0) return fail(MSOSM_ERR_IO, std::string("saver_create: io_uring_setup: ") + strerror(-r));
        }
    }
    *out = s;
    return MSOSM_OK;
}

extern "C" int msosm_saver_start(msosm_saver *s)
{
    if (!s) return set_error(MSOSM_ERR_ARG, "saver_start: null saver");
    if (s->started) return set_error(MSOSM_ERR_STATE, "saver_start: already started");
    s->started = true;
    s->own_threads = true;
    s->poller = std::thread(poll_loop, s);
    const bool uring = (s->flags & MSOSM_SAVE_URING) != 0;
    for (int k = 0; k < s->nslots; ++k) s->writers.emplace_back(writer_loop, s, k, uring);
    return MSOSM_OK;
}

// poll_cuda_state (data_save/msosm_data_save_uring.cpp:115-154) on the calling thread: hands each committed
// slot to its writer once its copy is done, then marks every slot done; returns when all save_count
// pushes are through. Every slot's writer must run meanwhile (msosm_saver_run_writer on other threads).
extern "C" int msosm_saver_run_poller(msosm_saver *s)
{
    if (!s) return set_error(MSOSM_ERR_ARG, "saver_run_poller: null saver");
    if (s->own_threads) return set_error(MSOSM_ERR_STATE, "saver_run_poller: msosm_saver_start runs the loops");
    s->started = true;
    poll_loop(s);
    return MSOSM_OK;
}

// save_to_disk / save_to_disk_pwrite (:156-258) of one slot on the calling thread, with io_uring (the saver
// made with MSOSM_SAVE_URING) or pwrite; returns once the poller has marked the slot done. Write errors are
// reported by msosm_saver_join.
extern "C" int msosm_saver_run_writer(msosm_saver *s, int slot, int use_uring)
{
    if (!s) return set_error(MSOSM_ERR_ARG, "saver_run_writer: null saver");
    if (slot < 0 || slot >= s->nslots) return set_error(MSOSM_ERR_ARG, "saver_run_writer: slot index out of range");
    if (s->own_threads) return set_error(MSOSM_ERR_STATE, "saver_run_writer: msosm_saver_start runs the loops");
    if (use_uring && !(s->flags & MSOSM_SAVE_URING))
        return set_error(MSOSM_ERR_STATE, "saver_run_writer: io_uring writes need a saver made with MSOSM_SAVE_URING");
    s->started = true;
    writer_loop(s, slot, use_uring != 0);
    return MSOSM_OK;
}

extern "C" int msosm_saver_acquire(msosm_saver *s, void **slot_data)
{
    if (!s || !slot_data) return set_error(MSOSM_ERR_ARG, "saver_acquire: null argument");
    if (s->committed >= s->save_count) return set_error(MSOSM_ERR_STATE, "saver_acquire: all pushes already saved");
    const int k = s->fill_index;
    if (!s->started.load() && s->state[k].load(std::memory_order_acquire) != EMPTY)
        return set_error(MSOSM_ERR_STATE, "saver_acquire: every slot is full and saving has not started");
    unsigned spins = 0;
    while (s->state[k].load(std::memory_order_acquire) != EMPTY) cpu_relax(spins);
    *slot_data = s->slots + (size_t)k * s->slot_bytes;
    return MSOSM_OK;
}

static int commit_slot(msosm_saver *s, void *stream, bool on_stream)
{
    const int k = s->fill_index;
    if (s->state[k].load(std::memory_order_acquire) != EMPTY)
        return set_error(MSOSM_ERR_STATE, "saver_commit: the slot was not acquired");
    if (on_stream) {
        if (!s->pinned) return set_error(MSOSM_ERR_STATE, "saver_commit_stream: slots are not pinned device-visible memory");
        const hipError_t e = hipEventRecord(s->events[k], static_cast<hipStream_t>(stream));
        if (e != hipSuccess) return set_error(MSOSM_ERR_HIP, std::string("saver_commit_stream: ") + hipGetErrorString(e));
    }
    s->has_event[k] = on_stream ? 1 : 0;
    s->bad[k] = 0;
    s->seq[k] = s->committed;
    s->committed += 1;
    s->state[k].store(COPYING, std::memory_order_release);
    s->fill_index = (k + 1) % s->nslots;
    return MSOSM_OK;
}

extern "C" int msosm_saver_commit_stream(msosm_saver *s, void *stream)
{
    if (!s) return set_error(MSOSM_ERR_ARG, "saver_commit_stream: null saver");
    return commit_slot(s, stream, true);
}

extern "C" int msosm_saver_commit_ready(msosm_saver *s)
{
    if (!s) return set_error(MSOSM_ERR_ARG, "saver_commit_ready: null saver");
    return commit_slot(s, nullptr, false);
}

extern "C" int msosm_saver_join(msosm_saver *s)
{
    if (!s) return set_error(MSOSM_ERR_ARG, "saver_join: null saver");
    if (!s->started) return set_error(MSOSM_ERR_STATE, "saver_join: not started");
    if (s->committed < s->save_count)
        return set_error(MSOSM_ERR_STATE, "saver_join: only " + std::to_string(s->committed) + " of " +
                                              std::to_string(s->save_count) + " pushes committed");
    if (s->poller.joinable()) s->poller.join();
    for (auto &t : s->writers)
        if (t.joinable()) t.join();
    s->writers.clear();
    int rc = MSOSM_OK;
    for (size_t i = 0; i < s->fds.size(); ++i) {
        if (s->fds[i] >= 0 && ::close(s->fds[i]) != 0 && s->err_code == MSOSM_OK)
            record_failure(s, MSOSM_ERR_IO, "saver_join: close " + s->paths[i] + ": " + strerror(errno));
        s->fds[i] = -1;
    }
    if (s->err_code != MSOSM_OK) rc = set_error(s->err_code, s->err);
    return rc;
}

extern "C" int msosm_saver_destroy(msosm_saver *s)
{
    if (!s) return MSOSM_OK;
    if (s->own_threads && (s->poller.joinable() || !s->writers.empty())) {
        // abandoned mid-run: release the threads by marking every slot done
        s->save_count.store(s->committed, std::memory_order_release);
        if (s->poller.joinable()) s->poller.join();
        for (auto &t : s->writers)
            if (t.joinable()) t.join();
    }
    saver_free(s);
    return MSOSM_OK;
}
