"""Precision stress (low band, high DM): 325-375 MHz, DMs 1500 (i+1)/128 for i = 0, 63, 127, FFT length
2^21 (M = 2^20, delaycount 700: the whole record lies inside the dispersion window, compared with zero
history on both sides). Variants: chirp phase in FP64 (this build's default) or in the reference's
FP32 form; output complex64 or FP16 pairs with a power-of-two block scale. Errors against the FP64
oracle are printed per DM and variant: rel-L2 and max |error| / max |y|.

Bars: FP64 phase + complex64 within rel-L2 1e-5; FP16 pairs within 2^-11 (FP16 keeps 11 bits); the
FP32 phase is reported (it is the reference's legacy form) and must be measurably worse than FP64.
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import make_input, oracle_design, rel_l2

pytestmark = pytest.mark.gpu

BW, F0 = 50e6, 325e6


def _maxrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_low_band_high_dm_precision_per_variant():
    dms = np.array([1500.0 * (i + 1) / 128 for i in (0, 63, 127)], dtype=np.float32)
    batch, calls = 2, 4
    res = {}
    for phase in (ms.PHASE_FP64, ms.PHASE_FP32):
        design = ms.Design(BW, F0, dms, phase_mode=phase)
        M = design.M
        assert M == 1 << 20 and design.delaycount == 700
        per = batch * M
        n = per * calls
        if not res:
            raw, x = make_input(ms.IN_C64, n, BW, F0, 750.0, 0.033, seed=44)
            delays, H = oracle_design(BW, F0, design.dms, M)
            print(f"case precision: n={n} M={M} dc={design.delaycount}; FP64 oracle ...", flush=True)
            ref = orc.msosm(x, M, delays, H)
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64 | ms.OUT_F16)
        c64, f16 = [], []
        for c in range(calls):
            plan.push(raw[c * per:(c + 1) * per])
            c64.append(plan.output_c64())
            f16.append(plan.output_f16()[0])
        plan.close()
        res[phase] = (np.concatenate(c64, axis=1), np.concatenate(f16, axis=1))
    for d, dm in enumerate(dms):
        line = []
        for phase, pname in ((ms.PHASE_FP64, "fp64-phase"), (ms.PHASE_FP32, "fp32-phase")):
            for k, oname in ((0, "c64"), (1, "fp16")):
                y = res[phase][k][d]
                line.append(f"{pname}/{oname} rel-L2 {rel_l2(y, ref[d]):.2e} max {_maxrel(y, ref[d]):.2e}")
        print(f"  DM {dm:.2f}: " + "; ".join(line), flush=True)
        assert rel_l2(res[ms.PHASE_FP64][0][d], ref[d]) < 1e-5
        assert rel_l2(res[ms.PHASE_FP64][1][d], ref[d]) < 2.0 ** -11
    # the reference's FP32 phase loses accuracy at this band and DM range
    worst32 = max(rel_l2(res[ms.PHASE_FP32][0][d], ref[d]) for d in range(len(dms)))
    worst64 = max(rel_l2(res[ms.PHASE_FP64][0][d], ref[d]) for d in range(len(dms)))
    assert worst32 > 10 * worst64
