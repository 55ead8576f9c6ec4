"""GPU parity of the accumulate dedispersion form (FFT lengths 2^14, 2^15, 2^16) against the FP64 oracle.

The accumulate launch splits each unit's inverse transform by decimation in time into R = L / 2^14
transforms of 2^14 points that stay in the CU, accumulating the kept outputs in registers. It is a
different factorisation from the two-pass forms, so its bar is the oracle's (rel-L2 1e-5, fold
counts bit-exact) and bit-identity with itself under every schedule and launch form.
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import check_u16_codes, make_input, oracle_design, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
# (band, DMs) giving each length by the reference's minimum-order rule
BANDS = {14: (16e6, 1392e6, [5.0, 25.0, 50.0]), 15: (32e6, 1384e6, [20.0, 60.0, 120.0]),
         16: (64e6, 1218e6, [10.0, 50.0, 100.0])}
ACC = ms.ORDER_XCD_DM_MAJOR | ms.SCHEDULE_ACCUM


def _design(log2L):
    bw, f0, dms = BANDS[log2L]
    d = ms.Design(bw, f0, np.asarray(dms, dtype=np.float32))
    assert 2 * d.M == 1 << log2L, (log2L, d.M)
    return bw, f0, d


def _run(design, raw, fmt, batch, out, wg=0, order=ACC, mode=ms.LAUNCH_PERSISTENT, fold=None):
    per = batch * design.M
    plan = ms.Plan.from_design(design, batch, in_format=fmt, out_flags=out)
    plan.set_schedule(wg, order)
    plan.set_launch_mode(mode)
    if fold is not None:
        plan.attach_fold(fold)
    outs = []
    for c in range(raw.shape[0] // per):
        plan.push(raw[c * per:(c + 1) * per])
        if out & ms.OUT_C64:
            outs.append(plan.output_c64())
        elif out & ms.OUT_U16:
            outs.append(plan.output_u16())
        elif out & ms.OUT_F16:
            outs.append(plan.output_f16()[0])
    plan.synchronize()
    plan.close()
    return np.concatenate(outs, axis=1) if outs else None


@pytest.mark.parametrize("log2L", [14, 15, 16])
def test_accum_series_and_fold_match_oracle(log2L):
    bw, f0, design = _design(log2L)
    batch, calls, nbins, period = 4, 3, 512, 0.0011
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_I8, 1 << (n - 1).bit_length(), bw, f0, float(design.dms[-1]), period, seed=50 + log2L)
    raw, x = raw[:n], x[:n]
    print(f"case accumulate L=2^{log2L} dms={list(design.dms)} dc={design.delaycount}", flush=True)
    fold = ms.Fold(period, bw, per, nbins, num_series=len(design.dms))
    got = _run(design, raw, ms.IN_I8, batch, ms.OUT_C64, fold=fold)
    _, sums, cnts = fold.read()
    fold.close()
    delays, H = oracle_design(bw, f0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    for d in range(len(design.dms)):
        err = rel_l2(got[d], ref[d])
        rs, rc = np.zeros(nbins), np.zeros(nbins, dtype=np.uint64)
        t0 = 0.0
        for c in range(calls):
            orc.fold(np.abs(ref[d, c * per:(c + 1) * per]) ** 2, nbins, t0, 1.0 / bw, period, rs, rc)
            t0 += float(per) * (1.0 / bw)
        srel = float(np.max(np.abs(sums[d] - rs) / rs))
        print(f"  DM {design.dms[d]}: rel-L2 {err:.3e}, fold counts equal {np.array_equal(cnts[d], rc)}, "
              f"sum rel err {srel:.2e}", flush=True)
        assert err < TOL
        assert np.array_equal(cnts[d], rc)
        assert srel < TOL


@pytest.mark.parametrize("log2L", [14, 15, 16])
def test_accum_uint16_and_fp16_outputs(log2L):
    bw, f0, design = _design(log2L)
    batch, calls = 2, 2
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_U16, 1 << (n - 1).bit_length(), bw, f0, float(design.dms[-1]), 0.0011, seed=60 + log2L)
    raw, x = raw[:n], x[:n]
    print(f"case accumulate outputs L=2^{log2L}", flush=True)
    delays, H = oracle_design(bw, f0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    codes = _run(design, raw, ms.IN_U16, batch, ms.OUT_U16)
    frac, worst = check_u16_codes(codes, ref)
    print(f"  uint16: max code diff {worst}, frac differing (all at ties) {frac:.2e}", flush=True)
    f16 = _run(design, raw, ms.IN_U16, batch, ms.OUT_F16)
    for d in range(len(design.dms)):
        err = rel_l2(f16[d], ref[d])
        print(f"  DM {design.dms[d]}: FP16 rel-L2 {err:.3e}", flush=True)
        assert err < 2.0 ** -11


@pytest.mark.parametrize("log2L", [14, 15, 16])
def test_accum_bit_identical_across_schedules_and_launch_forms(log2L):
    # 40 trials x 4 blocks: units chained through workgroups, every unit order, half the CUs, the
    # one-launch-per-DM form; all give the same bits
    bw, f0, _ = _design(log2L)
    dms = np.linspace(0.0, BANDS[log2L][2][-1], 40).astype(np.float32)
    design = ms.Design(bw, f0, dms, M=1 << (log2L - 1))
    batch, calls = 4, 2
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), bw, f0, float(dms[-1]), 0.0011, seed=70 + log2L)
    raw = np.ascontiguousarray(raw[:n])
    ref = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64)
    cases = [(0, o | ms.SCHEDULE_ACCUM, ms.LAUNCH_PERSISTENT) for o in range(7)]
    cases += [(128, ms.ORDER_XCD_TILE16 | ms.SCHEDULE_ACCUM, ms.LAUNCH_PERSISTENT),
              (8, ms.ORDER_LINEAR | ms.SCHEDULE_ACCUM, ms.LAUNCH_PERSISTENT),
              (0, ms.ORDER_XCD_DM_MAJOR | ms.SCHEDULE_ACCUM | ms.SCHEDULE_SPLIT, ms.LAUNCH_PERSISTENT),
              (0, ACC, ms.LAUNCH_PER_DM)]
    for wg, order, mode in cases:
        print(f"case accumulate L=2^{log2L} workgroups={wg} order={order} mode={mode}", flush=True)
        got = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64, wg=wg, order=order, mode=mode)
        assert np.array_equal(got, ref)
    # the two-pass form of the same plan layout agrees to rounding
    seq = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64, order=ms.ORDER_XCD_DM_MAJOR)
    assert rel_l2(seq, ref) < 2e-6


def test_accum_tangent_fold_matches_oracle():
    bw, f0 = 16e6, 1392e6
    dms = np.array([122.9, 123.456, 124.0], dtype=np.float32)
    design = ms.Design(bw, f0, dms)
    assert 2 * design.M == 1 << 14
    batch, calls, nbins, period = 4, 3, 128, 0.0011
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), bw, f0, 123.456, period, seed=17)
    raw, x = raw[:n], x[:n]
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_NONE)
    plan.set_schedule(0, ACC)
    for i, dm in enumerate(dms):
        plan.set_dm_tangent(i, ms.design_transfer_function_ddm(bw, f0, float(dm), design.M))
    fold = ms.Fold(period, bw, per, nbins, num_series=len(dms))
    tfold = ms.Fold(period, bw, per, nbins, num_series=len(dms))
    plan.attach_fold(fold)
    for c in range(calls):
        plan.push_tangent(raw[c * per:(c + 1) * per], tfold)
    plan.synchronize()
    prof, _, cnts = fold.read()
    dprof, _, dcnts = tfold.read()
    plan.close()
    fold.close()
    tfold.close()
    assert np.array_equal(cnts, dcnts)
    delays, H = oracle_design(bw, f0, dms, design.M)
    dH = np.stack([orc.transfer_function_ddm(bw, f0, float(dm), design.M) for dm in dms])
    y, dy = orc.msosm(x, design.M, delays, H), orc.msosm(x, design.M, delays, dH)
    for d in range(len(dms)):
        p, dp, c = orc.profile_tangent(y[d], dy[d], nbins, 0.0, 1.0 / bw, period, call_len=per)
        assert np.array_equal(c, cnts[d])
        ep, edp = rel_l2(prof[d], p), rel_l2(dprof[d], dp)
        print(f"  DM {dms[d]}: profile rel-L2 {ep:.2e}, d profile/dDM rel-L2 {edp:.2e}", flush=True)
        assert ep < TOL and edp < 1e-4
