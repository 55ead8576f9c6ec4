"""Shared helpers for the parity tests: seeded inputs, GPU runs through the C ABI, oracle runs."""
from __future__ import annotations

import numpy as np

import msosm_amd as ms
from oracle import oracle as orc


def rel_l2(a, b) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def make_input(fmt: int, n: int, fs: float, f0: float, dm: float, period: float, seed: int = 7):
    """Seeded dispersed pulsar in the requested input format.

    Returns (raw samples as pushed to the GPU, the same samples as complex128 for the oracle)."""
    x = orc.pulsar(n, fs, f0, dm, period, duty=0.03, amp=0.5, sigma=1.0, seed=seed)
    if fmt == ms.IN_C64:
        raw = x.astype(np.complex64)
        return raw, raw.astype(np.complex128)
    if fmt == ms.IN_I8:
        q = orc.quantize_int8(x)
        return q.reshape(-1, 2), orc.int8_to_complex(q)
    if fmt == ms.IN_U16:
        s = 2000.0 / np.sqrt(np.mean(np.abs(x) ** 2) / 2)
        re = np.clip(np.floor(x.real * s + 32768.0), 0, 65535).astype(np.uint32)
        im = np.clip(np.floor(x.imag * s + 32768.0), 0, 65535).astype(np.uint32)
        packed = (im << 16) | re
        return packed, orc.unpack_u16(packed)
    raise ValueError(fmt)


def oracle_design(bw, f0, dms, M):
    delays = np.stack([orc.segmentation(bw, f0, float(np.float32(dm)), M)[0] for dm in dms])
    H = np.stack([orc.transfer_function(bw, f0, float(np.float32(dm)), M) for dm in dms])
    return delays, H


def gpu_series(design: ms.Design, raw: np.ndarray, fmt: int, batch: int, out=ms.OUT_C64):
    """Push raw through a plan call by call; returns the concatenated [D][N] output."""
    M = design.M
    per = batch * M
    assert raw.shape[0] % per == 0
    plan = ms.Plan.from_design(design, batch, in_format=fmt, out_flags=out)
    outs = []
    for c in range(raw.shape[0] // per):
        plan.push(raw[c * per:(c + 1) * per])
        outs.append(plan.output_c64() if out & ms.OUT_C64 else plan.output_u16())
    plan.close()
    return np.concatenate(outs, axis=1)


def check_u16_codes(got: np.ndarray, ref: np.ndarray, tie_window: float = 0.02):
    """uint16 codes from the GPU against the FP64 series: every code equals the oracle's encoding of the
    FP64 value, except where that value lies within tie_window of a rounding boundary (x + 32768 at
    k + 1/2), where FP32 rounding may legitimately pick the neighbour; those differ by one code.
    Returns (fraction of components that differ, max code difference)."""
    enc = np.vectorize(orc.encode_u16_pair, otypes=[np.uint32])
    ref_codes = enc(ref.real.astype(np.float32), ref.imag.astype(np.float32))
    worst, ndiff, ntot = 0, 0, 0
    for shift, comp in ((0, ref.real), (16, ref.imag)):
        g = ((got >> shift) & 0xFFFF).astype(np.int64)
        r = ((ref_codes >> shift) & 0xFFFF).astype(np.int64)
        d = np.abs(g - r)
        worst = max(worst, int(d.max()))
        v = comp + np.where(comp != 0, 32768.0, 0.0)
        near_tie = np.abs(v - np.floor(v) - 0.5) < tie_window
        assert np.all(near_tie[d > 0]), "a code differs away from a rounding boundary"
        ndiff += int((d > 0).sum())
        ntot += d.size
    assert worst <= 1
    return ndiff / ntot, worst


def write_dada(path, pol1: np.ndarray, pol2: np.ndarray | None, bw_mhz: float, fc_mhz: float):
    """A DADA file as psrdada/psrdada.cpp reads it: a 4096-byte ASCII header, then frames of 2048
    uint16-pair samples, one frame per polarisation in turn."""
    npol = 1 if pol2 is None else 2
    hdr = (f"HDR_VERSION 1.0\nHDR_SIZE    4096\nBW          {bw_mhz}\nFREQ        {fc_mhz}\nNBIT        16\n"
           f"NCHAN       1\nNDIM        2\nNPOL        {npol}\n").encode()
    frames = pol1.reshape(-1, 2048).astype(np.uint32)
    if npol == 2:
        frames = np.stack([frames, pol2.reshape(-1, 2048).astype(np.uint32)], axis=1).reshape(-1, 2048)
    with open(path, "wb") as fh:
        fh.write(hdr.ljust(4096, b"\0"))
        fh.write(frames.tobytes())


class _DevArray:
    """A device buffer of the library (plain pointer) as torch reads it in place."""

    def __init__(self, ptr: int, n: int, typestr: str):
        self.__cuda_array_interface__ = {"data": (int(ptr), False), "shape": (int(n),), "typestr": typestr,
                                         "version": 2}


def device_view(ptr: int, n: int, typestr: str = "<u4"):
    """torch tensor over n elements at a device pointer owned by the library (no copy)."""
    import torch

    return torch.as_tensor(_DevArray(ptr, n, typestr), device="cuda")
