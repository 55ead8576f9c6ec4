"""Host-side parity (no GPU): the library's design tables against the oracle, bit-exact for the
integer results (Nd, M, delaycount, per-bin delays, segments) and to complex64 rounding for H; plus
the C ABI's export list and argument validation."""
import ctypes
import os
import re

import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

GRIDS = {
    # name: (bw, f0, DM list, M or None for the reference rule)
    "cfg1_osm": (16e6, 1392e6, [50.0], None),
    "cfg2": (64e6, 1218e6, np.linspace(0, 100, 512), None),
    "cfg3_some": (32e6, 784e6, np.linspace(0, 1000, 1024)[::97], None),
    "cfg4_some": (50e6, 325e6, [1500.0 * (i + 1) / 128 for i in (0, 31, 127)], None),
    "cfg5": (16e6, 1392e6, 123.456 + np.linspace(-0.5, 0.5, 64), None),
    "descending": (16e6, 1392e6, [80.0, 40.0, 10.0], None),
}


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "msosm_amd.h")) as fh:
        declared = set(re.findall(r"\b(msosm_[a-z0-9_]+)\s*\(", fh.read()))
    assert declared == set(ms.EXPORTS)
    lib = ctypes.CDLL(ms.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing


@pytest.mark.parametrize("name", list(GRIDS))
def test_design_tables_bit_exact(name):
    bw, f0, dms, _ = GRIDS[name]
    dms = np.asarray(dms, dtype=np.float32)
    for dm in dms:
        nd = orc.nd(bw, f0, float(dm))
        assert ms.design_nd(bw, f0, float(dm)) == nd
        assert ms.design_osm_order(nd) == orc.osm_min_order(nd)
        if nd > 0:
            assert ms.design_min_order(nd) == orc.min_order(nd)
    design = ms.Design(bw, f0, dms)
    assert design.M == max(orc.min_order(orc.nd(bw, f0, float(d))) for d in dms)
    for i, dm in enumerate(dms):
        ref, dc = orc.segmentation(bw, f0, float(dm), design.M)
        assert np.array_equal(design.delays[i], ref), f"DM {dm}"
        assert design.delaycount_values[i] == dc
        assert np.array_equal(ms.design_build_segments(design.delays[i]), orc.build_segments(ref))


@pytest.mark.parametrize("name", ["cfg1_osm", "cfg2", "cfg5"])
def test_transfer_function_rounding(name):
    bw, f0, dms, _ = GRIDS[name]
    dms = np.asarray(dms, dtype=np.float32)[:8]
    design = ms.Design(bw, f0, dms)
    for i, dm in enumerate(dms):
        ref = orc.transfer_function(bw, f0, float(dm), design.M)
        err = np.max(np.abs(design.H[i].astype(np.complex128) - ref))
        assert err <= 1.2e-7 * max(1.0, np.max(np.abs(ref))), f"DM {dm}: {err}"
    # the single-DM entry point equals the batched one
    h1 = ms.design_transfer_function(bw, f0, float(dms[-1]), design.M)
    assert np.array_equal(h1, design.H[len(dms) - 1])


def test_fp32_phase_mode_matches_oracle_legacy():
    bw, f0, dm, M = 64e6, 1218e6, 57.3, 32768
    got = ms.design_transfer_function(bw, f0, dm, M, ms.PHASE_FP32)
    ref = orc.transfer_function(bw, f0, dm, M, phase_fp32=True)
    assert np.max(np.abs(got - ref)) < 2e-7


def test_fftpoint_override_rule():
    # fftpoint below 2*M_min falls back to M_min (msosm_gpu_batch.cpp:33-41)
    d = ms.Design(16e6, 1392e6, [50.0], fftpoint=4096)
    assert d.M == d.M_min == 8192
    d = ms.Design(16e6, 1392e6, [50.0], fftpoint=1 << 16)
    assert d.M == 1 << 15


def test_c_abi_argument_errors():
    with pytest.raises(ms.MsosmError) as e:
        ms.design_segmentation(16e6, 1392e6, 50.0, 3000)
    assert e.value.code == ms.ERR_ARG and "power of two" in str(e.value)
    with pytest.raises(ms.MsosmError) as e:
        ms.Plan(M=1 << 30, num_dms=1, batch=1, delaycount=1)
    assert e.value.code == ms.ERR_UNSUPPORTED
    with pytest.raises(ms.MsosmError) as e:
        ms.Plan(M=4096, num_dms=0, batch=1, delaycount=1)
    assert e.value.code == ms.ERR_ARG
    with pytest.raises(ms.MsosmError) as e:
        ms.Plan(M=4096, num_dms=1, batch=1, delaycount=70000)
    assert e.value.code == ms.ERR_ARG
    with pytest.raises(ms.MsosmError) as e:
        ms.Plan(M=4096, num_dms=1, batch=1, delaycount=1, in_format=7)
    assert e.value.code == ms.ERR_ARG
    # unit positions of the launches are 32-bit: num_dms x batch stays below 2^29
    with pytest.raises(ms.MsosmError) as e:
        ms.Plan(M=4096, num_dms=1 << 15, batch=1 << 14, delaycount=1)
    assert e.value.code == ms.ERR_ARG and "2^29" in str(e.value)


def test_uneven_order_rule():
    # per DM: the reference's minimum order for that DM alone, clamped to FFT lengths 2^14..2^21
    for bw, f0, dm in [(32e6, 784e6, 0.0), (32e6, 784e6, 1.0), (32e6, 784e6, 100.0), (32e6, 784e6, 1000.0),
                       (50e6, 325e6, 1500.0)]:
        nd = orc.nd(bw, f0, dm)
        expect = min(1 << 20, max(1 << 13, orc.min_order(nd) if nd > 0 else 0))
        assert ms.design_uneven_order(nd) == expect
    d = ms.UnevenDesign(32e6, 784e6, np.linspace(0, 1000, 16))
    assert d.M.min() == 1 << 13 and d.M.max() == 1 << 17
    for i in range(16):
        ref, dc = orc.segmentation(32e6, 784e6, float(d.dms[i]), int(d.M[i]))
        assert np.array_equal(d.delays[i], ref) and d.delaycount[i] == dc
