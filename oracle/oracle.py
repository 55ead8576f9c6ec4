"""ctypes front-end of the CPU oracle (test infrastructure only).

Loads ``oracle/_build/liboracle.so`` (built by ``make -C oracle``) and wraps each function with
numpy arrays. Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this module, and only as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.orc_fft.argtypes = [dp, C.c_long, C.c_int]
        L.orc_nd.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_nd.restype = C.c_int
        L.orc_min_order.argtypes = [C.c_int]
        L.orc_osm_min_order.argtypes = [C.c_ulong]
        L.orc_osm_min_order.restype = C.c_long
        L.orc_delaycount.argtypes = [C.c_int, C.c_int]
        L.orc_segmentation.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, ip]
        L.orc_build_segments.argtypes = [ip, C.c_int, C.c_int, ip, C.c_int]
        L.orc_transfer_function.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, dp]
        L.orc_transfer_function_ddm.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, dp]
        L.orc_msosm.argtypes = [dp, C.c_long, C.c_int, C.c_int, ip, dp, dp]
        L.orc_fold.argtypes = [dp, C.c_long, C.c_ulong, C.c_double, C.c_double, C.c_double, dp,
                               C.POINTER(C.c_ulonglong), lp]
        L.orc_fold.restype = None
        L.orc_fold_bin.argtypes = [C.c_long, C.c_ulong, C.c_double, C.c_double, C.c_double]
        L.orc_fold_bin.restype = C.c_long
        L.orc_unpack_u16.argtypes = [C.POINTER(C.c_uint32), C.c_long, dp]
        L.orc_unpack_u16.restype = None
        L.orc_encode_u16_pair.argtypes = [C.c_float, C.c_float]
        L.orc_encode_u16_pair.restype = C.c_uint32
        L.orc_pulsar.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_uint64, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised FP64 DFT, sign -1 = e^{-i} (forward), +1 = e^{+i}."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    if lib().orc_fft(_dp(a.view(np.float64)), a.size, sign) != 0:
        raise ValueError("fft: size must be a power of two")
    return a


def nd(bw: float, f0: float, dm: float) -> int:
    return int(lib().orc_nd(bw, f0, dm))


def min_order(nd_: int) -> int:
    return int(lib().orc_min_order(nd_))


def osm_min_order(nd_: int) -> int:
    return int(lib().orc_osm_min_order(nd_))


def delaycount(nd_: int, M: int) -> int:
    return int(lib().orc_delaycount(nd_, M))


def segmentation(bw: float, f0: float, dm: float, M: int):
    """Returns (delay_points[2M] int32, delaycount)."""
    d = np.empty(2 * M, dtype=np.int32)
    dc = lib().orc_segmentation(bw, f0, dm, M, _ip(d))
    if dc < 0:
        raise ValueError("segmentation failed")
    return d, int(dc)


def build_segments(delay: np.ndarray, max_len: int = 256) -> np.ndarray:
    d = np.ascontiguousarray(delay, dtype=np.int32)
    cap = d.size + 1
    seg = np.empty(3 * cap, dtype=np.int32)
    n = lib().orc_build_segments(_ip(d), d.size, max_len, _ip(seg), cap)
    if n < 0:
        raise ValueError("build_segments overflow")
    return seg[: 3 * n].reshape(n, 3)


def transfer_function(bw: float, f0: float, dm: float, M: int, phase_fp32: bool = False) -> np.ndarray:
    H = np.empty(2 * M, dtype=np.complex128)
    if lib().orc_transfer_function(bw, f0, dm, M, int(phase_fp32), _dp(H.view(np.float64))) != 0:
        raise ValueError("transfer_function failed")
    return H


def transfer_function_ddm(bw: float, f0: float, dm: float, M: int, phase_fp32: bool = False) -> np.ndarray:
    dH = np.empty(2 * M, dtype=np.complex128)
    if lib().orc_transfer_function_ddm(bw, f0, dm, M, int(phase_fp32), _dp(dH.view(np.float64))) != 0:
        raise ValueError("transfer_function_ddm failed")
    return dH


def msosm(x: np.ndarray, M: int, delays: np.ndarray, H: np.ndarray) -> np.ndarray:
    """FP64 MS-OSM over the whole record; delays [D][2M], H [D][2M] -> out [D][N]."""
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    dl = np.ascontiguousarray(delays, dtype=np.int32)
    HH = np.ascontiguousarray(H, dtype=np.complex128)
    D = dl.shape[0]
    assert dl.shape == (D, 2 * M) and HH.shape == (D, 2 * M)
    out = np.empty((D, xx.size), dtype=np.complex128)
    rc = lib().orc_msosm(_dp(xx.view(np.float64)), xx.size, M, D, _ip(dl), _dp(HH.view(np.float64)),
                         _dp(out.view(np.float64)))
    if rc != 0:
        raise ValueError(f"msosm failed ({rc})")
    return out


def fold(intensity: np.ndarray, nbins: int, t_start: float, dt: float, period: float,
         sums=None, counts=None, want_bins: bool = False):
    """Reference fold of one series; accumulates into (sums, counts) and returns them (+ bins)."""
    I = np.ascontiguousarray(intensity, dtype=np.float64)
    if sums is None:
        sums = np.zeros(nbins, dtype=np.float64)
    if counts is None:
        counts = np.zeros(nbins, dtype=np.uint64)
    bins = np.empty(I.size, dtype=np.int64) if want_bins else None
    lib().orc_fold(_dp(I), I.size, nbins, t_start, dt, period, _dp(sums),
                   counts.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                   bins.ctypes.data_as(C.POINTER(C.c_long)) if want_bins else None)
    return (sums, counts, bins) if want_bins else (sums, counts)


def profile_tangent(y: np.ndarray, dy: np.ndarray, nbins: int, t_start: float, dt: float, period: float,
                    call_len: int | None = None):
    """Folded profile p = sum(|y|^2)/count per bin and its derivative dp = sum(2 Re(conj(y) dy))/count.

    y is one dedispersed series and dy its derivative with respect to DM (the MS-OSM output is linear
    in H, so dy is the MS-OSM of the same input with dH/dDM). Samples with |y|^2 == 0 are skipped in
    both, as foldDataPhase_kernel (gpu/gpu_kernel.cu:462-496) skips them; counts do not depend on DM.
    With call_len, the series is folded in calls of call_len samples whose clock advances by
    call_len * dt per call, as Fold_GPU::fold_data_phase does (folding/fold_gpu.cpp:48-54)."""
    I = np.abs(y) ** 2
    J = 2.0 * (y.real * dy.real + y.imag * dy.imag)
    step = I.size if call_len is None else call_len
    sums = np.zeros(nbins, dtype=np.float64)
    counts = np.zeros(nbins, dtype=np.uint64)
    dsums = np.zeros(nbins, dtype=np.float64)
    t0 = t_start
    for s0 in range(0, I.size, step):
        seg = slice(s0, min(s0 + step, I.size))
        _, _, bins = fold(I[seg], nbins, t0, dt, period, sums, counts, want_bins=True)
        keep = I[seg] != 0.0
        np.add.at(dsums, bins[keep], J[seg][keep])
        t0 += float(step) * dt
    with np.errstate(invalid="ignore", divide="ignore"):
        p = np.where(counts > 0, sums / np.maximum(counts, 1), 0.0)
        dp = np.where(counts > 0, dsums / np.maximum(counts, 1), 0.0)
    return p, dp, counts


def fold_bin(i: int, nbins: int, t_start: float, dt: float, period: float) -> int:
    return int(lib().orc_fold_bin(i, nbins, t_start, dt, period))


def unpack_u16(packed: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(packed, dtype=np.uint32)
    out = np.empty(p.size, dtype=np.complex128)
    lib().orc_unpack_u16(p.ctypes.data_as(C.POINTER(C.c_uint32)), p.size, _dp(out.view(np.float64)))
    return out


def encode_u16_pair(re: float, im: float) -> int:
    return int(lib().orc_encode_u16_pair(re, im))


def pulsar(n: int, fs: float, f0: float, dm: float, period: float, duty: float = 0.03, amp: float = 0.5,
           sigma: float = 1.0, seed: int = 20261203) -> np.ndarray:
    out = np.empty(n, dtype=np.complex128)
    if lib().orc_pulsar(n, fs, f0, dm, period, duty, amp, sigma, seed, _dp(out.view(np.float64))) != 0:
        raise ValueError("pulsar: n must be a power of two")
    return out


def quantize_int8(x: np.ndarray, rms: float = 32.0) -> np.ndarray:
    """complex -> interleaved int8 I/Q: scale to the given rms, round half-even, clip to [-127,127]."""
    s = rms / np.sqrt(np.mean(np.abs(x) ** 2) / 2.0)
    re = np.clip(np.rint(x.real * s), -127, 127).astype(np.int8)
    im = np.clip(np.rint(x.imag * s), -127, 127).astype(np.int8)
    out = np.empty(2 * x.size, dtype=np.int8)
    out[0::2], out[1::2] = re, im
    return out


def int8_to_complex(q: np.ndarray) -> np.ndarray:
    q = q.astype(np.float64)
    return q[0::2] + 1j * q[1::2]


# ---- FP32 CPU port (the timed CPU baseline; oracle/cpu_msosm.c) ----
_CPU_PATH = os.path.join(_HERE, "_build", "liboracle_cpu.so")
_cpu = None


def cpu_lib():
    global _cpu
    if _cpu is None:
        if not os.path.exists(_CPU_PATH):
            build()
        L = C.CDLL(_CPU_PATH)
        L.cpu_msosm_backend.argtypes = [C.c_int]
        L.cpu_msosm_backend.restype = C.c_char_p
        L.cpu_msosm_fold.argtypes = [C.c_void_p, C.c_long, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_ulong,
                                     C.c_double, C.c_double, C.c_double, C.c_long, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int, C.c_int]
        L.cpu_msosm_fold.restype = C.c_int
        _cpu = L
    return _cpu


def cpu_backend(force_own: bool = False) -> str:
    """Name of the transform backend of the CPU port (MKL's FFTW3 interface, or its own FFT)."""
    return cpu_lib().cpu_msosm_backend(int(force_own)).decode()


def cpu_msosm_fold(x: np.ndarray, M: int, delays: np.ndarray, H: np.ndarray, nbins: int, t_start: float, dt: float,
                   period: float, nthreads: int = 0, force_own: bool = False, want_out: bool = False,
                   call_size: int = 0):
    """FP32 MS-OSM + fold of every DM on the CPU (OpenMP over DMs). x: complex baseband (any complex
    dtype, cast to complex64), delays: [D][2M] block delays, H: [D][2M] transfer functions (without
    the 1/(2M) of the inverse transform). The fold bins the record as calls of call_size samples
    (0: one call), t_start advancing by call_size*dt per call as Fold_GPU does. Returns (sums
    [D][nbins], counts [D][nbins], series [D][n] complex64 or None)."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    delays = np.ascontiguousarray(delays, dtype=np.uint16)
    H = np.ascontiguousarray(H, dtype=np.complex64)
    D, L = delays.shape
    if L != 2 * M or H.shape != (D, L):
        raise ValueError("delays and H must be [D][2M]")
    sums = np.zeros((D, nbins))
    cnts = np.zeros((D, nbins), dtype=np.uint64)
    out = np.zeros((D, x.size), dtype=np.complex64) if want_out else None
    rc = cpu_lib().cpu_msosm_fold(x.ctypes.data, x.size, M, D, delays.ctypes.data, H.ctypes.data, nbins, t_start, dt,
                                  period, call_size, sums.ctypes.data, cnts.ctypes.data,
                                  out.ctypes.data if out is not None else None, nthreads, int(force_own))
    if rc != 0:
        raise RuntimeError(f"cpu_msosm_fold failed ({rc})")
    return sums, cnts, out
