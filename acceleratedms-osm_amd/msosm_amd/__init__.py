"""Python front-end of the MI355X MS-OSM library (ctypes over the C ABI in include/msosm_amd.h).

The compute path is the HIP library ``_lib/libmsosm_amd.so``; nothing here computes samples on the
CPU. Importing this module does not touch the GPU; creating a :class:`Plan` does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import NamedTuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_HERE, "_lib", "libmsosm_amd.so")

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_ALLOC, ERR_IO = range(7)
SAVE_URING, SAVE_DIRECT, SAVE_PINNED = 1, 2, 4
IN_U16, IN_I8, IN_C64 = 0, 1, 2
OUT_NONE, OUT_C64, OUT_U16, OUT_F16 = 0, 1, 2, 4
PHASE_FP64, PHASE_FP32 = 0, 1
IN_BYTES = {IN_U16: 4, IN_I8: 2, IN_C64: 8}
INPUT_HOST_ASYNC, INPUT_DEVICE, INPUT_HOST_SYNC = 0, 1, 2

# every symbol declared in include/msosm_amd.h
EXPORTS = [
    "msosm_last_error", "msosm_version",
    "msosm_design_nd", "msosm_design_min_order", "msosm_design_osm_order", "msosm_design_delaycount",
    "msosm_design_segmentation", "msosm_design_build_segments", "msosm_design_transfer_function",
    "msosm_design_transfer_functions",
    "msosm_plan_create", "msosm_plan_destroy", "msosm_plan_set_dm", "msosm_plan_push", "msosm_plan_output_c64",
    "msosm_plan_output_u16", "msosm_plan_output_f16", "msosm_plan_get_output", "msosm_plan_synchronize", "msosm_plan_reset",
    "msosm_plan_get_stream", "msosm_plan_get_output_stream", "msosm_plan_get_info", "msosm_plan_profile", "msosm_plan_profile_read",
    "msosm_fold_create", "msosm_fold_destroy", "msosm_fold_accumulate", "msosm_fold_read", "msosm_fold_reset",
    "msosm_fold_get_nbins", "msosm_plan_attach_fold", "msosm_plan_set_launch_mode",
    "msosm_design_uneven_order", "msosm_multi_create", "msosm_multi_destroy", "msosm_multi_get_info",
    "msosm_multi_set_dm", "msosm_multi_set_launch_mode", "msosm_multi_push", "msosm_multi_output_c64",
    "msosm_multi_get_output", "msosm_multi_fold", "msosm_multi_fold_read", "msosm_multi_synchronize",
    "msosm_multi_profile", "msosm_multi_profile_read",
    "msosm_design_transfer_function_ddm", "msosm_plan_set_dm_tangent", "msosm_plan_push_tangent",
    "msosm_multi_set_graph", "msosm_multi_autotune",
    "msosm_saver_create", "msosm_saver_start", "msosm_saver_acquire", "msosm_saver_commit_stream",
    "msosm_saver_commit_ready", "msosm_saver_join", "msosm_saver_destroy", "msosm_saver_file_name",
    "msosm_saver_run_poller", "msosm_saver_run_writer",
    "msosm_plan_set_schedule", "msosm_plan_get_schedule", "msosm_plan_autotune", "msosm_stream_wait",
    "msosm_simulate_params", "msosm_simulate_noise", "msosm_simulate_chirp", "msosm_simulate_pulsar_block",
    "msosm_simulate_params_new", "msosm_simulate_noise_new", "msosm_simulate_pulsar_new",
    "msosm_simulate_params_pulse", "msosm_simulate_pulsar",
    "msosm_host_alloc", "msosm_host_free", "msosm_fold_series_noalloc",
]
LAUNCH_PERSISTENT, LAUNCH_PER_DM, LAUNCH_PER_CLASS = 0, 1, 2
ORDER_XCD_DM_MAJOR, ORDER_LINEAR, ORDER_XCD_BLOCK_MAJOR, ORDER_XCD_TILE4, ORDER_XCD_TILE8 = 0, 1, 2, 3, 4
ORDER_XCD_TILE16, ORDER_XCD_TILE2 = 5, 6
SCHEDULE_SPLIT = 8  # flag on the order: role-split launch (FFT lengths 2^14 and 2^16, plain epilogue)
SCHEDULE_ACCUM = 16  # flag on the order: accumulate launch (FFT lengths 2^14..2^16, no scratch slot)


class MsosmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class PlanDesc(C.Structure):
    _fields_ = [("M", C.c_int), ("num_dms", C.c_int), ("batch", C.c_int), ("delaycount", C.c_int),
                ("in_format", C.c_int), ("out_flags", C.c_int), ("device", C.c_int), ("stream", C.c_void_p)]


class MultiDesc(C.Structure):
    _fields_ = [("num_dms", C.c_int), ("M", C.POINTER(C.c_int)), ("delaycount", C.POINTER(C.c_int)),
                ("push_blocks", C.c_int), ("in_format", C.c_int), ("out_flags", C.c_int), ("device", C.c_int),
                ("stream", C.c_void_p)]


class PlanInfo(C.Structure):
    _fields_ = [("L", C.c_int), ("FBS", C.c_int), ("workgroups", C.c_int), ("blocks_pushed", C.c_longlong),
                ("device_bytes", C.c_size_t)]


_lib = None


def build(jobs: int = 8) -> None:
    """Compile the HIP library in-tree (hipcc --offload-arch=gfx950)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)


def lib():
    """The loaded HIP library; raises if it has not been built (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C {PKG_DIR}")
    L = C.CDLL(LIB_PATH)
    ip, fp, vp, dp = C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_void_p, C.POINTER(C.c_double)
    sig = {
        "msosm_last_error": ([], C.c_char_p),
        "msosm_version": ([], C.c_int),
        "msosm_design_nd": ([C.c_float, C.c_float, C.c_float, ip], C.c_int),
        "msosm_design_min_order": ([C.c_int, ip], C.c_int),
        "msosm_design_osm_order": ([C.c_ulong, C.POINTER(C.c_ulong)], C.c_int),
        "msosm_design_delaycount": ([C.c_int, C.c_int, ip], C.c_int),
        "msosm_design_segmentation": ([C.c_float, C.c_float, C.c_float, C.c_int, ip, ip], C.c_int),
        "msosm_design_build_segments": ([ip, C.c_int, C.c_int, ip, C.c_int, ip], C.c_int),
        "msosm_design_transfer_function": ([C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, fp], C.c_int),
        "msosm_design_transfer_functions": ([C.c_float, C.c_float, fp, C.c_int, C.c_int, C.c_int, fp, C.c_int],
                                            C.c_int),
        "msosm_plan_create": ([C.POINTER(PlanDesc), C.POINTER(vp)], C.c_int),
        "msosm_plan_destroy": ([vp], C.c_int),
        "msosm_fold_series_noalloc": ([vp, vp, C.c_longlong, C.c_longlong, C.c_int, C.c_ulong, C.c_double,
                                       C.c_double, C.c_double, vp, vp, vp], C.c_int),
        "msosm_plan_set_dm": ([vp, C.c_int, ip, fp], C.c_int),
        "msosm_plan_push": ([vp, vp, C.c_int], C.c_int),
        "msosm_plan_output_c64": ([vp, C.POINTER(vp)], C.c_int),
        "msosm_plan_output_u16": ([vp, C.POINTER(vp)], C.c_int),
        "msosm_plan_output_f16": ([vp, C.POINTER(vp), C.POINTER(vp)], C.c_int),
        "msosm_plan_get_output": ([vp, C.c_int, vp], C.c_int),
        "msosm_plan_synchronize": ([vp], C.c_int),
        "msosm_plan_reset": ([vp], C.c_int),
        "msosm_plan_get_stream": ([vp, C.POINTER(vp)], C.c_int),
        "msosm_plan_get_output_stream": ([vp, C.POINTER(vp)], C.c_int),
        "msosm_plan_get_info": ([vp, C.POINTER(PlanInfo)], C.c_int),
        "msosm_plan_profile": ([vp, C.c_int], C.c_int),
        "msosm_plan_profile_read": ([vp, C.c_int, dp, C.POINTER(C.c_longlong)], C.c_int),
        "msosm_fold_create": ([C.c_double, C.c_double, C.c_ulong, C.c_ulong, C.c_int, C.c_int, C.POINTER(vp)],
                              C.c_int),
        "msosm_fold_destroy": ([vp], C.c_int),
        "msosm_fold_accumulate": ([vp, vp, vp, C.c_longlong, vp], C.c_int),
        "msosm_fold_read": ([vp, fp, dp, C.POINTER(C.c_ulonglong)], C.c_int),
        "msosm_fold_reset": ([vp], C.c_int),
        "msosm_fold_get_nbins": ([vp, C.POINTER(C.c_ulong)], C.c_int),
        "msosm_plan_attach_fold": ([vp, vp], C.c_int),
        "msosm_plan_set_launch_mode": ([vp, C.c_int], C.c_int),
        "msosm_design_uneven_order": ([C.c_int, ip], C.c_int),
        "msosm_multi_create": ([C.POINTER(MultiDesc), C.POINTER(vp)], C.c_int),
        "msosm_multi_destroy": ([vp], C.c_int),
        "msosm_multi_get_info": ([vp, C.c_int, ip, ip, C.POINTER(C.c_longlong)], C.c_int),
        "msosm_multi_set_dm": ([vp, C.c_int, ip, fp], C.c_int),
        "msosm_multi_set_launch_mode": ([vp, C.c_int], C.c_int),
        "msosm_multi_push": ([vp, vp, C.c_int], C.c_int),
        "msosm_multi_output_c64": ([vp, C.c_int, C.POINTER(vp)], C.c_int),
        "msosm_multi_get_output": ([vp, C.c_int, vp], C.c_int),
        "msosm_multi_fold": ([vp, C.c_double, C.c_double, C.c_ulong], C.c_int),
        "msosm_multi_fold_read": ([vp, fp, dp, C.POINTER(C.c_ulonglong)], C.c_int),
        "msosm_multi_synchronize": ([vp], C.c_int),
        "msosm_multi_profile": ([vp, C.c_int], C.c_int),
        "msosm_multi_profile_read": ([vp, C.c_int, dp, C.POINTER(C.c_longlong)], C.c_int),
        "msosm_design_transfer_function_ddm": ([C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, fp], C.c_int),
        "msosm_plan_set_dm_tangent": ([vp, C.c_int, fp], C.c_int),
        "msosm_plan_push_tangent": ([vp, vp, C.c_int, vp], C.c_int),
        "msosm_multi_set_graph": ([vp, C.c_int], C.c_int),
        "msosm_multi_autotune": ([vp, vp, C.c_int, ip, dp], C.c_int),
        "msosm_saver_create": ([C.POINTER(C.c_char_p), C.c_int, C.c_size_t, C.c_int, C.c_ulong, C.c_int,
                                C.POINTER(vp)], C.c_int),
        "msosm_saver_start": ([vp], C.c_int),
        "msosm_saver_run_poller": ([vp], C.c_int),
        "msosm_saver_run_writer": ([vp, C.c_int, C.c_int], C.c_int),
        "msosm_saver_acquire": ([vp, C.POINTER(vp)], C.c_int),
        "msosm_saver_commit_stream": ([vp, vp], C.c_int),
        "msosm_saver_commit_ready": ([vp], C.c_int),
        "msosm_saver_join": ([vp], C.c_int),
        "msosm_saver_destroy": ([vp], C.c_int),
        "msosm_saver_file_name": ([C.c_char_p, C.c_char_p, fp, C.c_int, C.c_int, C.c_char_p, C.c_size_t], C.c_int),
        "msosm_plan_set_schedule": ([vp, C.c_int, C.c_int], C.c_int),
        "msosm_plan_get_schedule": ([vp, ip, ip], C.c_int),
        "msosm_plan_autotune": ([vp, vp, C.c_int, ip, ip, dp], C.c_int),
        "msosm_stream_wait": ([vp, vp], C.c_int),
        "msosm_simulate_params": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong, C.c_ulong,
                                   C.POINTER(C.c_ulonglong)], C.c_int),
        "msosm_simulate_noise": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong, C.c_ulong, fp,
                                  C.c_ulonglong], C.c_int),
        "msosm_simulate_chirp": ([C.c_float, C.c_float, C.c_float, C.c_ulonglong, fp], C.c_int),
        "msosm_simulate_params_new": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong,
                                       C.POINTER(C.c_ulonglong)], C.c_int),
        "msosm_simulate_noise_new": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong, fp, C.c_ulonglong],
                                     C.c_int),
        "msosm_simulate_pulsar_new": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong, C.c_int,
                                       C.POINTER(C.c_uint32), fp, C.c_ulonglong], C.c_int),
        "msosm_host_alloc": ([C.c_ulonglong, C.POINTER(C.c_void_p)], C.c_int),
        "msosm_host_free": ([C.c_void_p], C.c_int),
        "msosm_simulate_params_pulse": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong,
                                         C.POINTER(C.c_ulonglong)], C.c_int),
        "msosm_simulate_pulsar": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong, C.c_int, C.c_float, C.c_int,
                                   C.POINTER(C.c_uint32), fp, C.c_ulonglong], C.c_int),
        "msosm_simulate_pulsar_block": ([C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulong, C.c_ulong, C.c_int,
                                         C.POINTER(C.c_uint32), fp, C.c_ulonglong], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != OK:
        raise MsosmError(rc, lib().msosm_last_error().decode())


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ------------------------------------------------------------------------------------------------
# design
# ------------------------------------------------------------------------------------------------
def design_nd(bw: float, f0: float, dm: float) -> int:
    v = C.c_int()
    _check(lib().msosm_design_nd(bw, f0, dm, C.byref(v)))
    return v.value


def design_min_order(nd: int) -> int:
    v = C.c_int()
    _check(lib().msosm_design_min_order(nd, C.byref(v)))
    return v.value


def design_osm_order(nd: int) -> int:
    v = C.c_ulong()
    _check(lib().msosm_design_osm_order(nd, C.byref(v)))
    return v.value


def design_delaycount(nd: int, M: int) -> int:
    v = C.c_int()
    _check(lib().msosm_design_delaycount(nd, M, C.byref(v)))
    return v.value


def design_segmentation(bw: float, f0: float, dm: float, M: int):
    d = np.empty(2 * M, dtype=np.int32)
    dc = C.c_int()
    _check(lib().msosm_design_segmentation(bw, f0, dm, M, _ip(d), C.byref(dc)))
    return d, dc.value


def design_build_segments(delay: np.ndarray, max_len: int = 256) -> np.ndarray:
    d = np.ascontiguousarray(delay, dtype=np.int32)
    cap = d.size + 1
    seg = np.empty(3 * cap, dtype=np.int32)
    n = C.c_int()
    _check(lib().msosm_design_build_segments(_ip(d), d.size, max_len, _ip(seg), cap, C.byref(n)))
    return seg[: 3 * n.value].reshape(n.value, 3)


def design_transfer_function(bw: float, f0: float, dm: float, M: int, phase_mode: int = PHASE_FP64) -> np.ndarray:
    H = np.empty(2 * M, dtype=np.complex64)
    _check(lib().msosm_design_transfer_function(bw, f0, dm, M, phase_mode, _fp(H.view(np.float32))))
    return H


def design_transfer_function_ddm(bw: float, f0: float, dm: float, M: int, phase_mode: int = PHASE_FP64) -> np.ndarray:
    """dH/dDM (2M complex64): the analytic DM derivative of the transfer function."""
    dH = np.empty(2 * M, dtype=np.complex64)
    _check(lib().msosm_design_transfer_function_ddm(bw, f0, dm, M, phase_mode, _fp(dH.view(np.float32))))
    return dH


def design_transfer_functions(bw: float, f0: float, dms, M: int, phase_mode: int = PHASE_FP64,
                              nthreads: int = 0) -> np.ndarray:
    d = np.ascontiguousarray(dms, dtype=np.float32)
    H = np.empty((d.size, 2 * M), dtype=np.complex64)
    _check(lib().msosm_design_transfer_functions(bw, f0, _fp(d), d.size, M, phase_mode, _fp(H.view(np.float32)),
                                                 nthreads))
    return H


def design_uneven_order(nd: int) -> int:
    v = C.c_int()
    _check(lib().msosm_design_uneven_order(nd, C.byref(v)))
    return v.value


class UnevenDesign:
    """Per-DM design: every DM at its own order (the reference rule per DM, FFT length 2^14..2^21)."""

    def __init__(self, bw: float, f0: float, dms, phase_mode: int = PHASE_FP64):
        self.bw, self.f0 = float(bw), float(f0)
        self.dms = np.ascontiguousarray(dms, dtype=np.float32)
        self.nd = np.array([design_nd(bw, f0, float(dm)) for dm in self.dms], dtype=np.int64)
        self.M = np.array([design_uneven_order(int(n)) for n in self.nd], dtype=np.int64)
        self.delays, self.H, self.delaycount = [], [], np.zeros(len(self.dms), dtype=np.int64)
        for M in np.unique(self.M):
            idx = np.nonzero(self.M == M)[0]
            H = design_transfer_functions(bw, f0, self.dms[idx], int(M), phase_mode)
            for j, i in enumerate(idx):
                d, dc = design_segmentation(bw, f0, float(self.dms[i]), int(M))
                self.delaycount[i] = dc
                self._put(i, d, H[j])

    def _put(self, i, d, h):
        while len(self.delays) <= i:
            self.delays.append(None)
            self.H.append(None)
        self.delays[i], self.H[i] = d, h


class Design:
    """Design tables for a DM list at one M (Prepare_MSOSM, msosm_prepare.cpp:3-164,215-284)."""

    def __init__(self, bw: float, f0: float, dms, fftpoint: int = 0, phase_mode: int = PHASE_FP64,
                 M: int | None = None):
        self.bw, self.f0 = float(bw), float(f0)
        self.dms = np.ascontiguousarray(dms, dtype=np.float32)
        self.nd = np.array([design_nd(bw, f0, float(dm)) for dm in self.dms], dtype=np.int64)
        self.M_values = np.array([design_min_order(int(n)) for n in self.nd], dtype=np.int64)
        self.M_min = int(self.M_values.max())
        if M is None:
            M = self.M_min
            if fftpoint:
                M = max(fftpoint // 2, self.M_min)
        self.M = int(M)
        segs = [design_segmentation(bw, f0, float(dm), self.M) for dm in self.dms]
        self.delays = np.stack([s[0] for s in segs])
        self.delaycount_values = np.array([s[1] for s in segs], dtype=np.int64)
        self.delaycount = int(self.delaycount_values.max())
        self.H = design_transfer_functions(bw, f0, self.dms, self.M, phase_mode)


# ------------------------------------------------------------------------------------------------
# device objects
# ------------------------------------------------------------------------------------------------
class Fold:
    """Fold_GPU (folding/fold_gpu.cpp) over num_series series."""

    def __init__(self, period: float, fs: float, size: int, nbins: int, num_series: int = 1, device: int = -1):
        h = C.c_void_p()
        _check(lib().msosm_fold_create(period, fs, size, nbins, num_series, device, C.byref(h)))
        self.h = h
        self.num_series = num_series
        nb = C.c_ulong()
        _check(lib().msosm_fold_get_nbins(h, C.byref(nb)))
        self.nbins = nb.value

    def accumulate(self, pol1_dev: int, pol2_dev: int | None = None, stride: int = 0, stream: int | None = None):
        _check(lib().msosm_fold_accumulate(self.h, C.c_void_p(pol1_dev), C.c_void_p(pol2_dev) if pol2_dev else None,
                                           stride, C.c_void_p(stream) if stream else None))

    def read(self):
        n = self.nbins * self.num_series
        prof = np.empty(n, dtype=np.float32)
        sums = np.empty(n, dtype=np.float64)
        cnts = np.empty(n, dtype=np.uint64)
        _check(lib().msosm_fold_read(self.h, _fp(prof), sums.ctypes.data_as(C.POINTER(C.c_double)),
                                     cnts.ctypes.data_as(C.POINTER(C.c_ulonglong))))
        shape = (self.num_series, self.nbins)
        return prof.reshape(shape), sums.reshape(shape), cnts.reshape(shape)

    def reset(self):
        _check(lib().msosm_fold_reset(self.h))

    def close(self):
        if self.h:
            lib().msosm_fold_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """The MSOSM_GPU_BATCH device pipeline (msosm_gpu_batch.cpp:12-245) behind the C ABI."""

    def __init__(self, M: int, num_dms: int, batch: int, delaycount: int, in_format: int = IN_C64,
                 out_flags: int = OUT_C64, device: int = -1, stream: int | None = None):
        desc = PlanDesc(M, num_dms, batch, delaycount, in_format, out_flags, device, stream)
        h = C.c_void_p()
        _check(lib().msosm_plan_create(C.byref(desc), C.byref(h)))
        self.h = h
        self.M, self.D, self.B, self.in_format, self.out_flags = M, num_dms, batch, in_format, out_flags
        self.fold = None

    @classmethod
    def from_design(cls, design: Design, batch: int, in_format: int = IN_C64, out_flags: int = OUT_C64, **kw):
        p = cls(design.M, len(design.dms), batch, design.delaycount, in_format, out_flags, **kw)
        for i in range(len(design.dms)):
            p.set_dm(i, design.delays[i], design.H[i])
        return p

    def set_dm(self, i: int, delay_points: np.ndarray, H: np.ndarray):
        d = np.ascontiguousarray(delay_points, dtype=np.int32)
        h = np.ascontiguousarray(H, dtype=np.complex64)
        assert d.size == 2 * self.M and h.size == 2 * self.M
        _check(lib().msosm_plan_set_dm(self.h, i, _ip(d), _fp(h.view(np.float32))))

    def push(self, x) -> None:
        """Push batch*M samples: a numpy array (host) or an int device address."""
        if isinstance(x, int):
            _check(lib().msosm_plan_push(self.h, C.c_void_p(x), INPUT_DEVICE))
            return
        a = np.ascontiguousarray(x)
        if a.nbytes != self.B * self.M * IN_BYTES[self.in_format]:
            raise ValueError(f"push: expected {self.B * self.M} samples, got {a.nbytes} bytes")
        _check(lib().msosm_plan_push(self.h, C.c_void_p(a.ctypes.data), INPUT_HOST_SYNC))

    def output_c64(self) -> np.ndarray:
        out = np.empty((self.D, self.B * self.M), dtype=np.complex64)
        _check(lib().msosm_plan_get_output(self.h, OUT_C64, C.c_void_p(out.ctypes.data)))
        self.synchronize()
        return out

    def output_u16(self) -> np.ndarray:
        out = np.empty((self.D, self.B * self.M), dtype=np.uint32)
        _check(lib().msosm_plan_get_output(self.h, OUT_U16, C.c_void_p(out.ctypes.data)))
        self.synchronize()
        return out

    def output_f16(self) -> tuple[np.ndarray, np.ndarray]:
        """FP16 pairs [D][B*M] as complex64 values (half * scale), and the [D][B] scales."""
        n = self.D * self.B * self.M
        raw = np.empty(n + self.D * self.B, dtype=np.uint32)
        _check(lib().msosm_plan_get_output(self.h, OUT_F16, C.c_void_p(raw.ctypes.data)))
        self.synchronize()
        halves = raw[:n].view(np.float16).astype(np.float32).reshape(self.D, self.B, self.M, 2)
        scales = raw[n:].view(np.float32).reshape(self.D, self.B)
        vals = (halves[..., 0] + 1j * halves[..., 1]) * scales[:, :, None]
        return vals.reshape(self.D, self.B * self.M).astype(np.complex64), scales

    def device_output_c64(self) -> int:
        p = C.c_void_p()
        _check(lib().msosm_plan_output_c64(self.h, C.byref(p)))
        return p.value

    def attach_fold(self, fold: Fold | None):
        _check(lib().msosm_plan_attach_fold(self.h, fold.h if fold else None))
        self.fold = fold

    def set_schedule(self, workgroups: int = 0, order: int = 0) -> None:
        _check(lib().msosm_plan_set_schedule(self.h, workgroups, order))

    def schedule(self) -> tuple[int, int]:
        w, o = C.c_int(), C.c_int()
        _check(lib().msosm_plan_get_schedule(self.h, C.byref(w), C.byref(o)))
        return w.value, o.value

    def autotune(self, input_dev: int, pushes: int = 3) -> tuple[int, int, float]:
        """Time every schedule on `pushes` pushes of device input; apply the fastest (resets the plan)."""
        w, o, t = C.c_int(), C.c_int(), C.c_double()
        _check(lib().msosm_plan_autotune(self.h, C.c_void_p(input_dev), pushes, C.byref(w), C.byref(o), C.byref(t)))
        return w.value, o.value, t.value

    def set_dm_tangent(self, i: int, dH: np.ndarray) -> None:
        """Upload dH/dDM of DM i for tangent pushes."""
        h = np.ascontiguousarray(dH, dtype=np.complex64)
        if h.size != 2 * self.M:
            raise ValueError("set_dm_tangent: dH must have 2M entries")
        _check(lib().msosm_plan_set_dm_tangent(self.h, i, _fp(h.view(np.float32))))

    def push_tangent(self, x, tangent_fold: Fold) -> None:
        """Push batch*M samples, folding |y|^2 into the attached fold (if any) and d|y|^2/dDM into
        tangent_fold."""
        if isinstance(x, int):
            _check(lib().msosm_plan_push_tangent(self.h, C.c_void_p(x), INPUT_DEVICE, tangent_fold.h))
            return
        a = np.ascontiguousarray(x)
        if a.nbytes != self.B * self.M * IN_BYTES[self.in_format]:
            raise ValueError(f"push_tangent: expected {self.B * self.M} samples, got {a.nbytes} bytes")
        _check(lib().msosm_plan_push_tangent(self.h, C.c_void_p(a.ctypes.data), INPUT_HOST_SYNC, tangent_fold.h))

    def synchronize(self):
        _check(lib().msosm_plan_synchronize(self.h))

    def reset(self):
        _check(lib().msosm_plan_reset(self.h))

    def stream(self) -> int:
        s = C.c_void_p()
        _check(lib().msosm_plan_get_stream(self.h, C.byref(s)))
        return s.value or 0

    def info(self) -> PlanInfo:
        i = PlanInfo()
        _check(lib().msosm_plan_get_info(self.h, C.byref(i)))
        return i

    def profile(self, enable: bool = True):
        _check(lib().msosm_plan_profile(self.h, int(enable)))

    def set_launch_mode(self, mode: int):
        _check(lib().msosm_plan_set_launch_mode(self.h, mode))

    def profile_read(self, kernel: int):
        ms, n = C.c_double(), C.c_longlong()
        _check(lib().msosm_plan_profile_read(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.h:
            lib().msosm_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Multi:
    """Uneven DM grid: one plan per length class on a shared stream (msosm_multi_* in the C ABI)."""

    def __init__(self, design: UnevenDesign, push_blocks: int, in_format: int = IN_C64, out_flags: int = OUT_C64,
                 device: int = -1):
        D = len(design.dms)
        self._M = np.ascontiguousarray(design.M, dtype=np.int32)
        self._dc = np.ascontiguousarray(design.delaycount, dtype=np.int32)
        desc = MultiDesc(D, _ip(self._M), _ip(self._dc), push_blocks, in_format, out_flags, device, None)
        h = C.c_void_p()
        _check(lib().msosm_multi_create(C.byref(desc), C.byref(h)))
        self.h = h
        self.D, self.in_format = D, in_format
        nc, ps = C.c_int(), C.c_longlong()
        _check(lib().msosm_multi_get_info(h, 0, None, C.byref(nc), C.byref(ps)))
        self.num_classes, self.push_samples = nc.value, ps.value
        for i in range(D):
            d = np.ascontiguousarray(design.delays[i], dtype=np.int32)
            hh = np.ascontiguousarray(design.H[i], dtype=np.complex64)
            _check(lib().msosm_multi_set_dm(h, i, _ip(d), _fp(hh.view(np.float32))))
        self.nbins = 0

    def set_launch_mode(self, mode: int):
        _check(lib().msosm_multi_set_launch_mode(self.h, mode))

    def set_graph(self, enable: bool = True):
        _check(lib().msosm_multi_set_graph(self.h, int(enable)))

    def autotune(self, input_dev: int, pushes: int = 3) -> tuple[int, float]:
        """Tune every length class's schedule, then keep the faster of one launch over all classes and
        one tuned launch per class (msosm_multi_autotune). Returns (launch mode, ms per push); resets
        the stream."""
        mode, t = C.c_int(), C.c_double()
        _check(lib().msosm_multi_autotune(self.h, C.c_void_p(input_dev), pushes, C.byref(mode), C.byref(t)))
        return mode.value, t.value

    def push(self, x) -> None:
        if isinstance(x, int):
            _check(lib().msosm_multi_push(self.h, C.c_void_p(x), INPUT_DEVICE))
            return
        a = np.ascontiguousarray(x)
        if a.nbytes != self.push_samples * IN_BYTES[self.in_format]:
            raise ValueError("push: wrong input size")
        _check(lib().msosm_multi_push(self.h, C.c_void_p(a.ctypes.data), INPUT_HOST_SYNC))

    def output(self, dm: int) -> np.ndarray:
        out = np.empty(self.push_samples, dtype=np.complex64)
        _check(lib().msosm_multi_get_output(self.h, dm, C.c_void_p(out.ctypes.data)))
        return out

    def fold(self, period: float, fs: float, nbins: int):
        _check(lib().msosm_multi_fold(self.h, period, fs, nbins))
        self.nbins = nbins

    def fold_read(self):
        n = self.D * self.nbins
        prof, sums, cnts = np.empty(n, np.float32), np.empty(n, np.float64), np.empty(n, np.uint64)
        _check(lib().msosm_multi_fold_read(self.h, _fp(prof), sums.ctypes.data_as(C.POINTER(C.c_double)),
                                           cnts.ctypes.data_as(C.POINTER(C.c_ulonglong))))
        sh = (self.D, self.nbins)
        return prof.reshape(sh), sums.reshape(sh), cnts.reshape(sh)

    def synchronize(self):
        _check(lib().msosm_multi_synchronize(self.h))

    def profile(self, enable: bool = True):
        _check(lib().msosm_multi_profile(self.h, int(enable)))

    def profile_read(self, kernel: int):
        ms, n = C.c_double(), C.c_longlong()
        _check(lib().msosm_multi_profile_read(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.h:
            lib().msosm_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def saver_file_name(directory: str, prefix: str, dms, i: int) -> str:
    """<dir>/<prefix>_DM<value>.dat with the reference's DM precision rule."""
    d = np.ascontiguousarray(dms, dtype=np.float32)
    buf = C.create_string_buffer(4096)
    _check(lib().msosm_saver_file_name(directory.encode(), prefix.encode(), _fp(d), d.size, i, buf, 4096))
    return buf.value.decode()


class Saver:
    """Output pipeline: slots of num_files * block_bytes, written to one file per series (C ABI
    msosm_saver_*; MSOSM_DataSave_Uring's slot protocol)."""

    def __init__(self, paths, block_bytes: int, slot_count: int, save_count: int, flags: int = SAVE_URING):
        arr = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
        h = C.c_void_p()
        _check(lib().msosm_saver_create(arr, len(paths), block_bytes, slot_count, save_count, flags, C.byref(h)))
        self.h, self.n, self.block = h, len(paths), block_bytes

    def start(self):
        _check(lib().msosm_saver_start(self.h))

    def run_poller(self):
        """poll_cuda_state on the calling thread (instead of start): returns when every push is through."""
        _check(lib().msosm_saver_run_poller(self.h))

    def run_writer(self, slot: int, uring: bool = True):
        """save_to_disk (uring) or save_to_disk_pwrite of one slot on the calling thread."""
        _check(lib().msosm_saver_run_writer(self.h, slot, int(uring)))

    def push_host(self, data: np.ndarray):
        """Copy one push (num_files * block_bytes bytes) into the next slot and commit it."""
        b = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        if b.size != self.n * self.block:
            raise ValueError("push_host: wrong size")
        p = C.c_void_p()
        _check(lib().msosm_saver_acquire(self.h, C.byref(p)))
        C.memmove(p, b.ctypes.data, b.size)
        _check(lib().msosm_saver_commit_ready(self.h))

    def acquire(self) -> int:
        p = C.c_void_p()
        _check(lib().msosm_saver_acquire(self.h, C.byref(p)))
        return p.value

    def commit_stream(self, stream: int | None):
        _check(lib().msosm_saver_commit_stream(self.h, C.c_void_p(stream or 0)))

    def join(self):
        _check(lib().msosm_saver_join(self.h))

    def close(self):
        if self.h:
            lib().msosm_saver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------------
# GPU pulsar simulator (simulated_complex_signal.cpp:74-277)
# ------------------------------------------------------------------------------------------------
class SimParams(NamedTuple):
    Np: int          # samples per period
    Nd: int          # dispersion samples
    range: int       # ceil(Nd / Np)
    fftpoint: int    # dispersion FFT length per chunk of 128 periods
    samples: int     # samples returned: (dummy_repeat + valid_repeat) * Np
    noise_len: int   # complex samples of the noise buffer
    start: int       # first extracted sample of a chunk's window
    chunks: int      # chunks of 128 periods


def simulate_params(bw: float, dm: float, f0: float, period: float, dummy_repeat: int = 0,
                    valid_repeat: int = 1) -> SimParams:
    v = (C.c_ulonglong * 8)()
    _check(lib().msosm_simulate_params(bw, dm, f0, period, dummy_repeat, valid_repeat, v))
    return SimParams(*[int(x) for x in v])


def simulate_noise(bw: float, dm: float, f0: float, period: float, dummy_repeat: int = 0,
                   valid_repeat: int = 1) -> np.ndarray:
    """The undispersed noise periods the simulator transforms (complex64, noise_len samples)."""
    p = simulate_params(bw, dm, f0, period, dummy_repeat, valid_repeat)
    out = np.empty(p.noise_len, dtype=np.complex64)
    _check(lib().msosm_simulate_noise(bw, dm, f0, period, dummy_repeat, valid_repeat, _fp(out.view(np.float32)),
                                      p.noise_len))
    return out


def simulate_chirp(bw: float, dm: float, f0: float, fftpoint: int) -> np.ndarray:
    out = np.empty(fftpoint, dtype=np.complex64)
    _check(lib().msosm_simulate_chirp(bw, dm, f0, fftpoint, _fp(out.view(np.float32))))
    return out


def simulate_pulsar_block(bw: float, dm: float, f0: float, period: float, dummy_repeat: int = 0,
                          valid_repeat: int = 1, device: int = 0, series: bool = False):
    """The dispersed record as uint16 pairs (uint32, real in the low half); with series=True also the
    complex64 series before the conversion."""
    p = simulate_params(bw, dm, f0, period, dummy_repeat, valid_repeat)
    codes = np.empty(p.samples, dtype=np.uint32)
    c64 = np.empty(p.samples, dtype=np.complex64) if series else None
    _check(lib().msosm_simulate_pulsar_block(bw, dm, f0, period, dummy_repeat, valid_repeat, device,
                                             codes.ctypes.data_as(C.POINTER(C.c_uint32)),
                                             _fp(c64.view(np.float32)) if series else None, p.samples))
    return (codes, c64) if series else codes


class SimParamsNew(NamedTuple):
    Np: int             # samples per period
    Nd: int             # dispersion samples
    range: int          # ceil(Nd / Np)
    fftpoint: int       # record FFT length
    samples: int        # samples returned: repeat * Np
    filter_point: int   # filter design length
    periods: int        # periods generated: repeat + 2 (range + 2)
    start: int          # first extracted sample: (range + 2) Np + Nd


def simulate_params_new(bw: float, dm: float, f0: float, period: float, repeat: int = 1) -> SimParamsNew:
    """Sizes of generate_pulsar_signal_new (simulated_complex_signal.cpp:279-467)."""
    v = (C.c_ulonglong * 8)()
    _check(lib().msosm_simulate_params_new(bw, dm, f0, period, repeat, v))
    return SimParamsNew(*[int(x) for x in v])


def simulate_noise_new(bw: float, dm: float, f0: float, period: float, repeat: int = 1) -> np.ndarray:
    """The undispersed seed-1 record, zero-padded to fftpoint samples (complex64)."""
    p = simulate_params_new(bw, dm, f0, period, repeat)
    out = np.empty(p.fftpoint, dtype=np.complex64)
    _check(lib().msosm_simulate_noise_new(bw, dm, f0, period, repeat, _fp(out.view(np.float32)), p.fftpoint))
    return out


def simulate_pulsar_new(bw: float, dm: float, f0: float, period: float, repeat: int = 1, device: int = 0,
                        series: bool = False):
    """generate_pulsar_signal_new's record as uint16 pairs (uint32, real in the low half); with
    series=True also the complex64 series before the conversion."""
    p = simulate_params_new(bw, dm, f0, period, repeat)
    codes = np.empty(p.samples, dtype=np.uint32)
    c64 = np.empty(p.samples, dtype=np.complex64) if series else None
    _check(lib().msosm_simulate_pulsar_new(bw, dm, f0, period, repeat, device,
                                           codes.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           _fp(c64.view(np.float32)) if series else None, p.samples))
    return (codes, c64) if series else codes


class SimParamsPulse(NamedTuple):
    Np: int          # samples per period
    Nd: int          # dispersion samples
    range: int       # ceil(Nd / Np), in float
    order: int       # response length: next power of two >= range * Np
    fftpoint: int    # 2 * order
    samples: int     # samples returned: repeat * Np


def simulate_params_pulse(bw: float, dm: float, f0: float, period: float, repeat: int = 1) -> SimParamsPulse:
    """Sizes of generate_pulsar_signal (simulated_complex_signal.cpp:469-706)."""
    v = (C.c_ulonglong * 8)()
    _check(lib().msosm_simulate_params_pulse(bw, dm, f0, period, repeat, v))
    return SimParamsPulse(*[int(x) for x in v[:6]])


def simulate_pulsar(bw: float, dm: float, f0: float, period: float, repeat: int = 1, add_noise: bool = False,
                    snr: float = 0.0, device: int = 0, series: bool = False):
    """generate_pulsar_signal's record (a dispersed two-impulse pulse, one period repeated) as uint16
    pairs (uint32, real in the low half); with series=True also the complex64 series."""
    p = simulate_params_pulse(bw, dm, f0, period, repeat)
    codes = np.empty(p.samples, dtype=np.uint32)
    c64 = np.empty(p.samples, dtype=np.complex64) if series else None
    _check(lib().msosm_simulate_pulsar(bw, dm, f0, period, repeat, int(add_noise), snr, device,
                                       codes.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       _fp(c64.view(np.float32)) if series else None, p.samples))
    return (codes, c64) if series else codes


# ------------------------------------------------------------------------------------------------
# Reader of the saved series (python_check/check.py:5-53)
# ------------------------------------------------------------------------------------------------
def read_complex_bin(filename, start_index: int = 0, length=None) -> np.ndarray:
    """The complex samples of a saved <prefix>_DM<dm>.dat file, from sample `start_index`, `length`
    samples (None: to the end of the file). Each sample is one uint32: the real part in the low 16
    bits, the imaginary part in the high 16, offset-binary (code - 32768) with code 0 read as 0.
    Returns complex128, as the reference's reader does."""
    with open(filename, "rb") as f:
        f.seek(4 * int(start_index))
        raw = np.fromfile(f, dtype=np.uint32, count=-1 if length is None else int(length))

    def component(c):
        v = c.astype(np.float64)
        v[c != 0] -= 32768.0
        return v

    return component(raw & np.uint32(0xFFFF)) + 1j * component(raw >> np.uint32(16))

