/*
 * msosm_amd.h -- C ABI of the MI355X-native MS-OSM coherent-dedispersion path.
 *
 * Plain pointers and sizes only. Host C++ (the reference-compatible classes in
 * acceleratedms-osm_amd/csrc/reference_api/) and any FFI (ctypes, cgo, JNI) bind to these symbols.
 * Citations name the reference interface each entry point replaces, as path:line relative to the
 * reference repository root.
 *
 * Conventions
 *  - Every function returns an int status (MSOSM_OK = 0); msosm_last_error() describes the last
 *    failure of the calling thread.
 *  - Complex values are interleaved float pairs (re, im), layout-compatible with cufftComplex,
 *    hipFloatComplex and float2.
 *  - "_dev" pointers are device (HBM) addresses; other buffers are host memory.
 *  - A plan is externally synchronised: one host thread drives it; all of its GPU work is ordered
 *    on the plan's stream.
 */
#ifndef MSOSM_AMD_H
#define MSOSM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MSOSM_OK = 0,
    MSOSM_ERR_ARG = 1,         /* invalid argument */
    MSOSM_ERR_HIP = 2,         /* HIP runtime failure */
    MSOSM_ERR_STATE = 3,       /* call not valid in the object's current mode */
    MSOSM_ERR_UNSUPPORTED = 4, /* valid request this build does not implement (e.g. FFT length) */
    MSOSM_ERR_ALLOC = 5,       /* host or device allocation failed */
    MSOSM_ERR_IO = 6           /* file I/O failed (output pipeline) */
};

/* Input sample formats */
enum {
    MSOSM_IN_U16 = 0, /* uint16 pair, offset binary, code 0 reserved for 0 (utils/include/globals.h:22-26) */
    MSOSM_IN_I8 = 1,  /* int8 I/Q pair */
    MSOSM_IN_C64 = 2  /* complex64 */
};

/* Output series selection (bit flags) */
enum {
    MSOSM_OUT_NONE = 0,
    MSOSM_OUT_C64 = 1, /* complex64 [D][B][M] (fold mode buffer, msosm_gpu_batch.cpp:141-145) */
    MSOSM_OUT_U16 = 2, /* uint16-pair [D][B][M] (save mode buffer, msosm_gpu_batch.cpp:146-150) */
    MSOSM_OUT_F16 = 4  /* FP16 pair [D][B][M] (re low half, im high half) + one float scale per (DM, block):
                          sample = half * scale; the scale is the power of two at or below the block's rms
                          (no reference counterpart; 4 B per sample instead of 8) */
};

/* Transfer-function phase arithmetic */
enum {
    MSOSM_PHASE_FP64 = 0, /* msosm_prepare.cpp:107-164 */
    MSOSM_PHASE_FP32 = 1  /* legacy float-w form, msosm_prepare.cpp:166-213 */
};

const char *msosm_last_error(void);
int msosm_version(void);

/* ------------------------------------------------------------------------------------------------
 * Design (host). Bit-identical to the reference's host arithmetic for Nd, M, delaycount,
 * segment boundaries and segments; the transfer function is computed in FP64 and rounded once.
 * --------------------------------------------------------------------------------------------- */

/* Prepare_MSOSM::Prepare_MSOSM Nd (dedispersion/msosm/msosm_prepare.cpp:12,33) */
int msosm_design_nd(float bw, float f0, float dm, int *nd);
/* Prepare_MSOSM::calculate_min_order, per-DM M_i (msosm_prepare.cpp:48-69) */
int msosm_design_min_order(int nd, int *M);
/* Prepare_OSM::calculate_min_order (dedispersion/overlap_save/osm_prepare.cpp:17-30) */
int msosm_design_osm_order(unsigned long nd, unsigned long *M);
/* delaycount of one DM (msosm_prepare.cpp:217); Nd == 0 (DM 0) gives 1 */
int msosm_design_delaycount(int nd, int M, int *delaycount);
/* Prepare_MSOSM::segmentation(i) (msosm_prepare.cpp:215-284): delay_points[2M] in blocks.
 * DM 0 yields all-zero delays (the reference leaves them uninitialised). */
int msosm_design_segmentation(float bw, float f0, float dm, int M, int *delay_points, int *delaycount);
/* MSOSM_GPU_BATCH::build_segments (dedispersion/msosm/msosm_gpu_batch.cpp:262-295):
 * (start, len, group) triples into seg[3*cap]. */
int msosm_design_build_segments(const int *delay_points, int fft_len, int max_len, int *seg, int cap, int *count);
/* Prepare_MSOSM::generate_dedisp_params(i) (msosm_prepare.cpp:107-164): H[2M] complex (2*2M floats). */
int msosm_design_transfer_function(float bw, float f0, float dm, int M, int phase_mode, float *H);
/* Same for num_dms DMs into H[num_dms][2M] using up to nthreads host threads (0 = all cores). */
int msosm_design_transfer_functions(float bw, float f0, const float *dms, int num_dms, int M, int phase_mode,
                                    float *H, int nthreads);
/* dH/dDM[2M] complex: the same design map applied to dc/dDM = -i (phi/DM) c, the analytic derivative of the
 * chirp of msosm_prepare.cpp:122-138 (no reference counterpart; it feeds the DM-gradient pass). */
int msosm_design_transfer_function_ddm(float bw, float f0, float dm, int M, int phase_mode, float *dH);

/* ------------------------------------------------------------------------------------------------
 * Device plan: the MSOSM_GPU_BATCH device pipeline (msosm_gpu_batch.cpp:12-151,178-245).
 * One push consumes batch*M new samples and produces batch*M output samples per DM.
 * --------------------------------------------------------------------------------------------- */
typedef struct msosm_plan msosm_plan;

typedef struct {
    int M;          /* half FFT length (fftpoint = 2M, a power of two in [2^12, 2^22]) */
    int num_dms;    /* DM trials D */
    int batch;      /* blocks of M samples per push (B); num_dms * batch below 2^29 */
    int delaycount; /* max over DMs of delaycount; spectra ring = roundup(delaycount + B - 1, B) slots */
    int in_format;  /* MSOSM_IN_* */
    int out_flags;  /* MSOSM_OUT_* */
    int device;     /* HIP device ordinal, -1 = current device */
    void *stream;   /* hipStream_t to run on; NULL = a stream owned by the plan */
} msosm_plan_desc;

typedef struct {
    int L, FBS, workgroups;
    long long blocks_pushed;
    size_t device_bytes;
} msosm_plan_info;

/* initialize_uint16 allocation half (msosm_gpu_batch.cpp:12-151); ring and history start at zero */
int msosm_plan_create(const msosm_plan_desc *desc, msosm_plan **plan);
/* frees device memory; never resets the device (cf. msosm_gpu_batch.cpp:257-260) */
int msosm_plan_destroy(msosm_plan *plan);
/* upload one DM's design: delay_points[2M] (values in [0, delaycount)) and H[2M] complex
 * (msosm_gpu_batch.cpp:54-93) */
int msosm_plan_set_dm(msosm_plan *plan, int dm_index, const int *delay_points, const float *H);
/* Where a pushed input lives */
enum {
    MSOSM_INPUT_HOST_ASYNC = 0, /* host memory copied asynchronously (pin it; keep it unchanged until the
                                   next synchronize), as the save-mode stream pipeline does */
    MSOSM_INPUT_DEVICE = 1,     /* device memory read in place; keep it valid until the next synchronize */
    MSOSM_INPUT_HOST_SYNC = 2   /* host memory copied before the call returns (fold-mode cudaMemcpy) */
};
/* filter_block_uint16 (msosm_gpu_batch.cpp:178-236) for any input format: batch*M new samples.
 * Host inputs go through two staging buffers. MSOSM_INPUT_HOST_ASYNC copies on the plan's copy-in
 * stream, ordered after the push two calls back has read the staging buffer, so the copy runs under the
 * previous push's kernels and this push's kernels wait for it (the reference's fft_stream, :203-209). */
int msosm_plan_push(msosm_plan *plan, const void *input, int input_mode);
/* get_output_pointer (dedispersion/include/msosm_gpu_batch.h:24): complex64 [D][B][M] device buffer */
int msosm_plan_output_c64(msosm_plan *plan, void **dev_ptr);
/* uint16-pair [D][B][M] device buffer (save mode) of the last push: the uint16 output has two buffers
 * that pushes alternate, so one push's copy-out overlaps the next push's kernels */
int msosm_plan_output_u16(msosm_plan *plan, void **dev_ptr);
/* FP16-pair [D][B][M] device buffer and its [D][B] float scales */
int msosm_plan_output_f16(msosm_plan *plan, void **dev_data, void **dev_scales);
/* get_output (msosm_gpu_batch.cpp:238-245): D2H copy of the last push's output, on the plan's copy-out
 * stream (the reference's output_stream) once that push's kernels are done; the push that next writes
 * the same device buffer waits for the copy. Asynchronous for pinned host memory: msosm_plan_synchronize
 * (or the copy-out stream) completes it. For MSOSM_OUT_F16, host_dst receives the [D][B][M] pairs
 * followed by the [D][B] float scales. */
int msosm_plan_get_output(msosm_plan *plan, int which, void *host_dst);
/* the stream msosm_plan_get_output copies on (MSOSM_DataSave_Uring records its slot events there) */
int msosm_plan_get_output_stream(msosm_plan *plan, void **stream);
/* synchronize (msosm_gpu_batch.cpp:247-250): the plan's compute and copy streams */
int msosm_plan_synchronize(msosm_plan *plan);
/* forget the stream history: zero ring + input history, block counter = 0 */
int msosm_plan_reset(msosm_plan *plan);
int msosm_plan_get_stream(msosm_plan *plan, void **stream);
int msosm_plan_get_info(msosm_plan *plan, msosm_plan_info *info);
/* How push launches the dedispersion kernel */
enum {
    MSOSM_LAUNCH_PERSISTENT = 0, /* one persistent launch over every (DM, block) unit (default) */
    MSOSM_LAUNCH_PER_DM = 1,     /* one launch per DM, as the reference's per-DM inverse loop
                                    (msosm_gpu_batch.cpp:196-199); kept as the comparison form */
    MSOSM_LAUNCH_PER_CLASS = 2   /* uneven grids only: one persistent launch per FFT length */
};
int msosm_plan_set_launch_mode(msosm_plan *plan, int mode);

/* Tile schedule of the persistent dedispersion launch (no reference counterpart): the number of
 * workgroups (0 = one per CU, rounded down to a multiple of 8; at most that) and the unit order.
 * A tile order whose tile does not divide the workgroups of an XCD (workgroups / 8) runs as the
 * XCD-grouped DM-major order, and every XCD-grouped order runs as the linear one when the workgroups
 * are not a multiple of 8. No schedule changes the output bits. */
enum {
    MSOSM_ORDER_XCD_DM_MAJOR = 0,   /* the workgroups of one XCD run consecutive blocks of one DM (default) */
    MSOSM_ORDER_LINEAR = 1,         /* workgroup w runs units w, w+G, ... in DM-major order */
    MSOSM_ORDER_XCD_BLOCK_MAJOR = 2, /* as the default, but consecutive units are one block of consecutive DMs */
    MSOSM_ORDER_XCD_TILE4 = 3,       /* an XCD's workgroups run 4 DMs x (workgroups/32) blocks at a time */
    MSOSM_ORDER_XCD_TILE8 = 4,       /* the same with 8 DMs */
    MSOSM_ORDER_XCD_TILE16 = 5,      /* the same with 16 DMs */
    MSOSM_ORDER_XCD_TILE2 = 6,       /* the same with 2 DMs */
    MSOSM_ORDER_MASK = 7,
    /* flag on the order: role-split launch, where half of each workgroup runs the first transform pass
     * of one unit while the other half runs the second pass of the previous unit on the same scratch
     * slot (FFT lengths 2^14 and 2^16, plain epilogue: complex64/uint16 output and/or fused fold;
     * other plans run the sequential form) */
    MSOSM_SCHEDULE_SPLIT = 8,
    /* flag on the order: accumulate launch (FFT lengths 2^14, 2^15 and 2^16, every epilogue), where a
     * unit's inverse transform is split by decimation in time into 1, 2 or 4 transforms of 2^14 points
     * that stay in the CU, the kept outputs accumulated in registers: no scratch slot round trip.
     * Lengths 2^17 and 2^18 (plans without FP16-pair output): each unit is split by decimation in
     * frequency into L / 2^16 sub-units of that form, each keeping every (L / 2^16)-th output, after a
     * gathering pass that writes the sub-units' spectra; a tangent push of such a plan runs the
     * sequential form. The unit order does not apply to the sub-unit launch (it runs a DM's blocks
     * and sub-units together on one XCD). Such plans hold the sub-unit spectra of one push in device
     * memory from creation: 8 * num_dms * batch * L bytes (17 GB for 507 trials x 16 blocks at
     * L = 2^18), and 4 * num_dms * batch bytes of arrival counters.
     * Takes precedence over MSOSM_SCHEDULE_SPLIT; other lengths run the sequential form */
    MSOSM_SCHEDULE_ACCUM = 16
};
int msosm_plan_set_schedule(msosm_plan *plan, int workgroups, int order);
int msosm_plan_get_schedule(msosm_plan *plan, int *workgroups, int *order);
/* Autotuner: times the dedispersion launch of every schedule candidate (unit orders x workgroup
 * counts {max, max/2}, each also with MSOSM_SCHEDULE_SPLIT and MSOSM_SCHEDULE_ACCUM where the plan
 * supports them) over
 * `pushes` pushes of device input (batch*M samples each, in the plan's input format; the first push
 * of each candidate is not timed), applies the fastest and returns it. The tuning pushes consume the
 * plan's state: the plan and its attached fold are reset afterwards. */
int msosm_plan_autotune(msosm_plan *plan, const void *input_dev, int pushes, int *best_workgroups,
                        int *best_order, double *best_ms);
/* HIP-event timing of the plan's kernels. kernel: 0 = forward FFT, 1 = dedispersion */
int msosm_plan_profile(msosm_plan *plan, int enable);
int msosm_plan_profile_read(msosm_plan *plan, int kernel, double *total_ms, long long *launches);

/* ------------------------------------------------------------------------------------------------
 * Fold: Fold_GPU (folding/fold_gpu.cpp:4-68) generalised to num_series series (one per DM).
 * --------------------------------------------------------------------------------------------- */
typedef struct msosm_fold msosm_fold;

/* Fold_GPU::Fold_GPU (fold_gpu.cpp:4-25): period [s], fs [Hz], size = samples per call,
 * nbins (0 = period*fs samples, as time_bin = -1) */
int msosm_fold_create(double period, double fs, unsigned long size, unsigned long nbins, int num_series, int device,
                      msosm_fold **fold);
int msosm_fold_destroy(msosm_fold *fold);
/* calculate_intensity + fold_data_phase (fold_gpu.cpp:27-54): series s at pol1_dev + s*series_stride
 * complex elements; pol2_dev != NULL selects the two-polarisation amplitude sqrt(|a|^2+|b|^2),
 * otherwise the power |a|^2. stream: hipStream_t or NULL (legacy default stream). */
/* work queued on waiting_stream after this call starts after the work queued so far on producer_stream
 * (an event dependency; the reference ran both polarisations on the default stream instead) */
int msosm_stream_wait(void *waiting_stream, void *producer_stream);
int msosm_fold_accumulate(msosm_fold *fold, const void *pol1_dev, const void *pol2_dev, long long series_stride,
                          void *stream);
/* get_folded_data (fold_gpu.cpp:56-60): profile[num_series][nbins] = sum/count (0 where empty);
 * any output pointer may be NULL */
int msosm_fold_read(msosm_fold *fold, float *profile, double *sums, unsigned long long *counts);

/* Fold without allocation. No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops or harnesses that forbid hipMalloc). One call does
 * what msosm_fold_create + msosm_fold_accumulate + msosm_fold_read's profile do for one call of `size`
 * samples starting at time t_start (dt = 1/fs): intensity of num_series series at pol1_dev +
 * s*series_stride complex64 elements (pol2_dev != NULL: two polarisations, as msosm_fold_accumulate),
 * FP64 sums and u64 counts per phase bin, then profile_dev[s][b] = sum/count (0 where empty) as float,
 * and, when counts_dev != NULL, the counts. All pointers are device pointers. The sums and counts are
 * kept in the library's static device array of MSOSM_FOLD_NOALLOC_MAX_ENTRIES bins (16 MiB, held on each
 * GPU the library runs on until the process ends), cleared by each call: num_series*nbins beyond that is
 * rejected (MSOSM_ERR_ARG). Kernels run on `stream` (NULL: the default stream); nothing is synchronised.
 * Every call on one GPU shares that array, so calls that overlap on one GPU (other streams or threads)
 * race. Calls no allocating routine; its first call on a device looks the array up
 * (hipGetSymbolAddress). */
#define MSOSM_FOLD_NOALLOC_MAX_ENTRIES (1u << 20)
int msosm_fold_series_noalloc(const void *pol1_dev, const void *pol2_dev, long long size, long long series_stride,
                              int num_series, unsigned long nbins, double t_start, double fs, double period,
                              float *profile_dev, unsigned long long *counts_dev, void *stream);
int msosm_fold_reset(msosm_fold *fold);
int msosm_fold_get_nbins(msosm_fold *fold, unsigned long *nbins);
/* fused fold: every following push folds the power of every DM's output inside the dedispersion
 * kernel (size must equal batch*M, num_series == num_dms, nbins <= 2048). NULL detaches. */
int msosm_plan_attach_fold(msosm_plan *plan, msosm_fold *fold);

/* DM-derivative pass (no reference counterpart; for gradient DM refinement). dH[2M] from
 * msosm_design_transfer_function_ddm for DM i. */
int msosm_plan_set_dm_tangent(msosm_plan *plan, int i, const float *dH);
/* A push that dedisperses with H (folding |y|^2 into the attached fold, if any) and with dH/dDM,
 * folding d|y|^2/dDM = 2 Re(conj(y) dy/dDM) into tangent_fold (same bins, size batch*M, one series
 * per DM; samples with |y|^2 == 0 are skipped as in the primal fold). tangent_fold's profile is then
 * d profile / d DM. No series outputs are written by this push. */
int msosm_plan_push_tangent(msosm_plan *plan, const void *input, int input_mode, msosm_fold *tangent_fold);

/* ------------------------------------------------------------------------------------------------
 * Uneven DM grids (no reference counterpart: the reference uses one M for every DM,
 * msosm_prepare.cpp:67). Each DM keeps its own M; DMs of equal M run as one plan ("length class");
 * all classes share one stream and one input stream, so a push dedisperses the same
 * push_blocks * max(M) samples for every DM.
 * --------------------------------------------------------------------------------------------- */
typedef struct msosm_multi msosm_multi;

typedef struct {
    int num_dms;
    const int *M;          /* per-DM half FFT length (power of two, FFT length 2^12..2^22) */
    const int *delaycount; /* per-DM delaycount */
    int push_blocks;       /* blocks of the largest M per push */
    int in_format;         /* MSOSM_IN_* */
    int out_flags;         /* MSOSM_OUT_* (MSOSM_OUT_C64 for per-DM series) */
    int device;            /* -1 = current */
    void *stream;          /* NULL = owned stream */
} msosm_multi_desc;

/* per-DM order: the reference's minimum order for this DM alone (msosm_prepare.cpp:59-61),
 * clamped to FFT lengths 2^14..2^21 */
int msosm_design_uneven_order(int nd, int *M);
int msosm_multi_create(const msosm_multi_desc *desc, msosm_multi **multi);
int msosm_multi_destroy(msosm_multi *multi);
int msosm_multi_get_info(msosm_multi *multi, int dm, int *M, int *num_classes, long long *push_samples);
/* delay_points and H hold 2*M[dm] entries */
int msosm_multi_set_dm(msosm_multi *multi, int dm, const int *delay_points, const float *H);
/* MSOSM_LAUNCH_PERSISTENT (default): one persistent launch over the units of every length class
 * (lengths 2^14..2^21, at most 8 classes, rings below 4 GiB each, fold bins within the LDS left by
 * the largest tile set; otherwise one persistent launch per class). MSOSM_LAUNCH_PER_CLASS: one
 * persistent launch per class. MSOSM_LAUNCH_PER_DM: one launch per DM. */
int msosm_multi_set_launch_mode(msosm_multi *multi, int mode);
/* enable != 0: in the one-launch form, run each push as one hipGraph (counter fill, per-class forward
 * transforms, history and fold bins, then the multi-length dedispersion), built on the first push and
 * re-parameterised on later ones. Profiling (kernel 1) then times the whole push graph. */
int msosm_multi_set_graph(msosm_multi *multi, int enable);
int msosm_multi_push(msosm_multi *multi, const void *input, int input_mode);
/* Autotuner of an uneven grid (no reference counterpart; BASELINE's per-length-class tuning): every
 * length class's plan tunes its tile schedule (msosm_plan_autotune on the same device input: unit
 * order, workgroups, role split / accumulate form where the length has them), then a push is timed as
 * one launch over all classes (MSOSM_LAUNCH_PERSISTENT) and as one tuned launch per class
 * (MSOSM_LAUNCH_PER_CLASS) over `pushes` pushes (the first not timed), and the faster becomes the launch
 * mode. Every class plan and fold is reset afterwards. best_ms: time per push of the kept mode. */
int msosm_multi_autotune(msosm_multi *multi, const void *input_dev, int pushes, int *best_mode, double *best_ms);
/* complex64 series of one DM for the last push (push_samples values) */
int msosm_multi_output_c64(msosm_multi *multi, int dm, void **dev_ptr);
int msosm_multi_get_output(msosm_multi *multi, int dm, void *host_dst);
/* fused fold of every DM into nbins phase bins (replaces any earlier fold) */
int msosm_multi_fold(msosm_multi *multi, double period, double fs, unsigned long nbins);
int msosm_multi_fold_read(msosm_multi *multi, float *profile, double *sums, unsigned long long *counts);
int msosm_multi_synchronize(msosm_multi *multi);
int msosm_multi_profile(msosm_multi *multi, int enable);
int msosm_multi_profile_read(msosm_multi *multi, int kernel, double *total_ms, long long *launches);

/* ------------------------------------------------------------------------------------------------
 * Host output pipeline: MSOSM_DataSave_Uring (data_save/msosm_data_save_uring.cpp:40-288).
 * slot_count slots of num_files * block_bytes bytes each; push k's block of series i goes to file i at
 * byte offset k * block_bytes. The producer acquires the next slot in turn, fills it (a device-to-host
 * copy on a stream, or host code), and commits it; a poll thread hands filled slots to one writer
 * thread per slot, which writes them with io_uring (raw system calls) or pwrite, resuming short
 * writes. An I/O error does not stop the producer or the other writes (every block is attempted);
 * msosm_saver_join reports the first one. The reference exits on the first error instead.
 * --------------------------------------------------------------------------------------------- */
typedef struct msosm_saver msosm_saver;
enum {
    MSOSM_SAVE_URING = 1,  /* io_uring writes (else pwrite, data_save/msosm_data_save_uring.cpp:224-258) */
    MSOSM_SAVE_DIRECT = 2, /* O_DIRECT files (:63); buffered where the file system, or the kernel for the
                              slot memory (pinned HIP host memory cannot be pinned for direct I/O), refuses it */
    MSOSM_SAVE_PINNED = 4  /* slots in pinned host memory (hipHostMalloc, :80): needed by commit_stream */
};
/* config_save + initialize_uring (:40-101); paths[num_files] name the files (see saver_file_name) */
int msosm_saver_create(const char *const *paths, int num_files, size_t block_bytes, int slot_count,
                       unsigned long save_count, int flags, msosm_saver **saver);
/* start_saving (:260-268): the poll thread and the writer threads */
int msosm_saver_start(msosm_saver *saver);
/* Or the caller's own threads run the loops instead of msosm_saver_start. poll_cuda_state (:115-154):
 * hands each committed slot to its writer once its copy has completed, then marks every slot done;
 * returns when all save_count pushes are through (every slot's writer must run meanwhile). */
int msosm_saver_run_poller(msosm_saver *saver);
/* save_to_disk (:156-222, use_uring != 0; the saver needs MSOSM_SAVE_URING) or save_to_disk_pwrite
 * (:224-258) of one slot; returns once the poller has marked the slot done. */
int msosm_saver_run_writer(msosm_saver *saver, int slot, int use_uring);
/* copy_to_slot, first half (:103-107): wait until the next slot is free; *slot_data = its buffer */
int msosm_saver_acquire(msosm_saver *saver, void **slot_data);
/* copy_to_slot, second half (:108-112): the acquired slot is filled once the work queued so far on
 * stream completes (an event is recorded on it); commit_ready: it is filled already */
int msosm_saver_commit_stream(msosm_saver *saver, void *stream);
int msosm_saver_commit_ready(msosm_saver *saver);
/* join_saving (:270-288): after save_count commits, wait for every write and close the files;
 * returns the first I/O error (MSOSM_ERR_IO) if any */
int msosm_saver_join(msosm_saver *saver);
int msosm_saver_destroy(msosm_saver *saver);
/* make_filename with the DM precision rule (:9-36,56-65): <dir>/<prefix>_DM<dm[i]>.dat */
int msosm_saver_file_name(const char *dir, const char *prefix, const float *dms, int num_dms, int i, char *buf,
                          size_t cap);

/* ---- page-locked host memory, for the reference classes' cudaMallocHost / cudaFreeHost buffers
 * (e.g. simulated_complex_signal.cpp:484-485, 491). bytes = 0 gives NULL. */
int msosm_host_alloc(unsigned long long bytes, void **ptr);
int msosm_host_free(void *ptr);

/* ---- GPU pulsar simulator: SimulatedComplexSignal::generate_pulsar_signal_block
 * (simulated_pulsar/simulated_complex_signal.cpp:74-277) with the dispersion injected on the GPU.
 * bw (= fs), dm, f0, period as the reference's float constructor arguments (:16-29);
 * dummy_repeat + valid_repeat periods of Np samples are returned. The noise periods (seeds 1 and 2)
 * and the chirp are built on the host as there; each chunk of 128 periods is dispersed with one
 * complex FFT of fftpoint points on the device. Deviations: nothing is cached in or read from files
 * (simulated_*.bin, random_reference.bin), and the noise buffer is zero-padded so that every chunk's
 * window lies inside it (the reference reads past its buffer unless repeat is a multiple of 128). */
#define MSOSM_SIM_NPARAMS 8
/* out[8] = {Np, Nd, range, fftpoint, samples returned, noise buffer length, first extracted sample
 * of a chunk's window, chunks} (:16-29, 80-116) */
int msosm_simulate_params(float bw, float dm, float f0, float period, unsigned long dummy_repeat,
                          unsigned long valid_repeat, unsigned long long *out);
/* the noise periods (generate_test_pulse, :43-72, 119-161) as interleaved complex64, len >= the
 * noise buffer length (complex samples) */
int msosm_simulate_noise(float bw, float dm, float f0, float period, unsigned long dummy_repeat,
                         unsigned long valid_repeat, float *periods, unsigned long long len);
/* the dispersion chirp e^{+i phi} at fftpoint bins (:166-198), interleaved complex64 */
int msosm_simulate_chirp(float bw, float dm, float f0, unsigned long long fftpoint, float *H);
/* the simulated record (:200-262): uint16 pairs (x + 32768 truncated as on x86, real in the low
 * half) and/or the complex64 series before the conversion, len >= repeat * Np; on `device`
 * (-1: the current device) */
int msosm_simulate_pulsar_block(float bw, float dm, float f0, float period, unsigned long dummy_repeat,
                                unsigned long valid_repeat, int device, uint32_t *out_u16, float *out_c64,
                                unsigned long long len);
/* SimulatedComplexSignal::generate_pulsar_signal_new (simulated_complex_signal.cpp:279-467): repeat
 * periods cut from one record of repeat + 2 (range + 2) periods drawn from one seed-1 stream, dispersed
 * by one FFT of fftpoint >= that record + Nd points through the causal response (the inverse FFT of the
 * chirp on filter_point >= Nd + 1 bins, its last Nd taps then tap 0), and read from
 * (range + 2) Np + Nd on. out[8] = {Np, Nd, range, fftpoint, samples returned, filter_point,
 * periods generated, first extracted sample} (:284-345, 431) */
int msosm_simulate_params_new(float bw, float dm, float f0, float period, unsigned long repeat,
                              unsigned long long *out);
/* the zero-padded record (generate_test_pulse over every period, :330-336), len >= fftpoint */
int msosm_simulate_noise_new(float bw, float dm, float f0, float period, unsigned long repeat, float *record,
                             unsigned long long len);
/* the simulated record as in msosm_simulate_pulsar_block, len >= repeat * Np. Deviation: filters whose
 * design length is below 2^12 or is 2^23 are designed with an FP64 inverse FFT on the host (the
 * reference's is FP32); the other transforms are FP32 on the device, as the reference's are FP32. */
int msosm_simulate_pulsar_new(float bw, float dm, float f0, float period, unsigned long repeat, int device,
                              uint32_t *out_u16, float *out_c64, unsigned long long len);
/* SimulatedComplexSignal::generate_pulsar_signal (simulated_complex_signal.cpp:469-706), as
 * msosm_save_simulated.cpp:115 and benchmark/test_msosm_speed.cpp:160 call it: impulses of 32767 at
 * Np/2 and Np/2 + 2 in each of range = ceil(Nd / Np) periods, dispersed through the response (the
 * inverse FFT of the chirp on order >= range Np bins, placed at [order, 2 order) of a 2 order FFT),
 * the first Np samples summed over those periods and repeated `repeat` times. out[8] = {Np, Nd, range,
 * order, 2 order, samples returned, 0, 0}. dm must be > 0 (at 0 the reference's order is undefined);
 * add_noise must be 0, as both callers pass: the reference's noise branch rewrites period 0 while
 * other threads copy it, so it returns MSOSM_ERR_UNSUPPORTED (snr is then unused). Deviation: the
 * response is designed with an FP64 inverse FFT on the host when order is below 2^12 or is 2^23. */
int msosm_simulate_params_pulse(float bw, float dm, float f0, float period, unsigned long repeat,
                                unsigned long long *out);
int msosm_simulate_pulsar(float bw, float dm, float f0, float period, unsigned long repeat, int add_noise, float snr,
                          int device, uint32_t *out_u16, float *out_c64, unsigned long long len);

#ifdef __cplusplus
}
#endif
#endif /* MSOSM_AMD_H */
