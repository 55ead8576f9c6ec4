"""GPU parity of the sub-unit accumulate form (FFT lengths 2^17 and 2^18) against the FP64 oracle.

At these lengths a unit's kept outputs do not fit a CU, so the accumulate launch splits each unit
(DM, block) by decimation in frequency into Q = L / 2^16 sub-units: a gathering pass writes the
spectra X_c[k'] = sum_s Y[k' + s 2^16] e^{+2 pi i c (k' + s 2^16) / L}, and sub-unit c dedisperses X_c
with a 2^16-point accumulate transform, keeping the outputs y[M + Q u + c]. Another factorisation
again, so its bar is the oracle's (rel-L2 1e-5, fold counts bit-exact), agreement with the two-pass
form to FP32 rounding, and bit-identity with itself under every workgroup count and launch form.
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import check_u16_codes, make_input, oracle_design, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
# band 784-816 MHz: the reference's minimum order gives FFT length 2^17 for DM 127..505 and 2^18 for
# DM 506..2020 (delaycount <= 128)
BW, F0 = 32e6, 784e6
DMS = {17: [150.0, 320.0, 500.0], 18: [520.0, 1000.0, 1500.0]}
ACC = ms.ORDER_XCD_DM_MAJOR | ms.SCHEDULE_ACCUM


def _design(log2L):
    d = ms.Design(BW, F0, np.asarray(DMS[log2L], dtype=np.float32))
    assert 2 * d.M == 1 << log2L and d.delaycount <= 256, (log2L, d.M, d.delaycount)
    return d


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
    plan.synchronize()
    plan.close()
    return np.concatenate(outs, axis=1) if outs else None


@pytest.mark.parametrize("log2L", [17, 18])
def test_sub_accum_series_and_fold_match_oracle(log2L):
    design = _design(log2L)
    batch, calls, nbins, period = 2, 3, 512, 0.0011
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, float(design.dms[1]), period, seed=80 + log2L)
    raw, x = raw[:n], x[:n]
    print(f"case sub-unit accumulate L=2^{log2L} dms={list(design.dms)} dc={design.delaycount}", flush=True)
    fold = ms.Fold(period, BW, per, nbins, num_series=len(design.dms))
    got = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64, fold=fold)
    _, sums, cnts = fold.read()
    fold.close()
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    for d in range(len(design.dms)):
        err = rel_l2(got[d], ref[d])
        rs, rc = np.zeros(nbins), np.zeros(nbins, dtype=np.uint64)
        t0 = 0.0
        for c in range(calls):
            orc.fold(np.abs(ref[d, c * per:(c + 1) * per]) ** 2, nbins, t0, 1.0 / BW, period, rs, rc)
            t0 += float(per) * (1.0 / BW)
        srel = float(np.max(np.abs(sums[d] - rs) / rs))
        print(f"  DM {design.dms[d]}: rel-L2 {err:.3e}, fold counts equal {np.array_equal(cnts[d], rc)}, "
              f"sum rel err {srel:.2e}", flush=True)
        assert err < TOL
        assert np.array_equal(cnts[d], rc)
        assert srel < TOL


def test_sub_accum_uint16_output_codes():
    design = _design(17)
    batch, calls = 2, 2
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_U16, 1 << (n - 1).bit_length(), BW, F0, 320.0, 0.0011, seed=91)
    raw, x = raw[:n], x[:n]
    print("case sub-unit accumulate uint16 L=2^17", flush=True)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    codes = _run(design, raw, ms.IN_U16, batch, ms.OUT_U16)
    frac, worst = check_u16_codes(codes, ref)
    print(f"  uint16: max code diff {worst}, frac differing (all at ties) {frac:.2e}", flush=True)


@pytest.mark.parametrize("log2L", [17, 18])
def test_sub_accum_bit_identical_across_workgroups_and_launch_forms(log2L):
    # 24 trials x 2 blocks x Q sub-units: items chained through few workgroups, on all CUs, on half, and
    # in the one-launch-per-DM form; all give the same bits, and the two-pass form agrees to rounding
    dms = np.linspace(DMS[log2L][0], DMS[log2L][-1], 24).astype(np.float32)
    design = ms.Design(BW, F0, dms, M=1 << (log2L - 1))
    batch, calls = 2, 2
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, float(dms[12]), 0.0011, seed=95 + log2L)
    raw = np.ascontiguousarray(raw[:n])
    ref = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64)
    for wg, order, mode in [(8, ACC, ms.LAUNCH_PERSISTENT), (128, ACC, ms.LAUNCH_PERSISTENT),
                            (0, ms.ORDER_LINEAR | ms.SCHEDULE_ACCUM, ms.LAUNCH_PERSISTENT),
                            (0, ACC, ms.LAUNCH_PER_DM)]:
        print(f"case sub-unit accumulate L=2^{log2L} workgroups={wg} order={order} mode={mode}", flush=True)
        got = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64, wg=wg, order=order, mode=mode)
        assert np.array_equal(got, ref)
    seq = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64, order=ms.ORDER_XCD_DM_MAJOR)
    err = rel_l2(seq, ref)
    print(f"  two-pass vs sub-unit accumulate rel-L2 {err:.2e}", flush=True)
    assert err < 2e-6


def test_sub_accum_schedule_falls_back_for_tangent_pushes():
    # the sub-unit form has no tangent epilogue: a tangent push under the accumulate schedule runs the
    # two-pass form, bit for bit as a plan scheduled for it
    design = _design(17)
    batch, calls, nbins, period = 2, 2, 128, 0.0011
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 320.0, period, seed=97)
    raw = np.ascontiguousarray(raw[:n])
    res = []
    for order in (ACC, ms.ORDER_XCD_DM_MAJOR):
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_NONE)
        plan.set_schedule(0, order)
        for i, dm in enumerate(design.dms):
            plan.set_dm_tangent(i, ms.design_transfer_function_ddm(BW, F0, float(dm), design.M))
        fold = ms.Fold(period, BW, per, nbins, num_series=len(design.dms))
        tfold = ms.Fold(period, BW, per, nbins, num_series=len(design.dms))
        plan.attach_fold(fold)
        for c in range(calls):
            plan.push_tangent(raw[c * per:(c + 1) * per], tfold)
        plan.synchronize()
        res.append((fold.read(), tfold.read()))
        plan.close()
        fold.close()
        tfold.close()
    for (pa, ta), (pb, tb) in [(res[0], res[1])]:
        for x, y in zip(pa + ta, pb + tb):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("log2L", [17, 18])
@pytest.mark.parametrize("batch", [1, 3])
def test_sub_accum_odd_batches_match_oracle_and_launch_forms(log2L, batch):
    # one and three blocks per push: the gathering pass's double-buffered rows end on either buffer, and
    # at Q = 2 its one-block-ahead ring request repeats the last block; the series against the oracle,
    # and the same bits from the persistent launch (sub-units meeting before their writes) and the
    # one-launch-per-DM form (grids that are no multiple of 8 Q do not meet)
    design = _design(log2L)
    calls = 3
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, float(design.dms[1]), 0.0011,
                        seed=120 + log2L + batch)
    raw, x = np.ascontiguousarray(raw[:n]), x[:n]
    print(f"case sub-unit accumulate L=2^{log2L} batch={batch} calls={calls}", flush=True)
    got = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64)
    per_dm = _run(design, raw, ms.IN_C64, batch, ms.OUT_C64, mode=ms.LAUNCH_PER_DM)
    assert np.array_equal(got, per_dm)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    for d in range(len(design.dms)):
        err = rel_l2(got[d], ref[d])
        print(f"  DM {design.dms[d]}: rel-L2 {err:.3e}", flush=True)
        assert err < TOL


@pytest.mark.parametrize("log2L", [17, 18])
def test_sub_accum_one_output_launch_matches_general_launch(log2L):
    # a plan with one output and no fold runs the sub-unit launch that loads each next spectrum into
    # registers ahead of its stores; with a fold attached the general launch runs: the series agree bit
    # for bit, for complex64 and for uint16 output
    design = _design(log2L)
    batch, calls = 2, 3
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, float(design.dms[2]), 0.0011, seed=140 + log2L)
    raw = np.ascontiguousarray(raw[:n])
    for out in (ms.OUT_C64, ms.OUT_U16):
        print(f"case sub-unit accumulate L=2^{log2L} out={out}: one-output launch vs general launch", flush=True)
        one = _run(design, raw, ms.IN_C64, batch, out)
        fold = ms.Fold(0.0011, BW, per, 64, num_series=len(design.dms))
        gen = _run(design, raw, ms.IN_C64, batch, out, fold=fold)
        fold.close()
        assert np.array_equal(one, gen)


@pytest.mark.parametrize("log2L", [17, 18])
def test_sub_accum_series_across_ring_wrap(log2L):
    # enough pushes that the input ring (roundup(delaycount + batch - 1, batch) blocks) wraps: the gathering
    # pass reads the history through the wrapped slots; the one-output launch's series of three trials at
    # the low end of the length class against the oracle, before and after the first wrap
    dms = {17: [130.0, 135.0, 140.0], 18: [506.0, 510.0, 515.0]}[log2L]
    design = ms.Design(BW, F0, np.asarray(dms, dtype=np.float32))
    assert 2 * design.M == 1 << log2L
    batch = 4
    per = batch * design.M
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
    plan.set_schedule(0, ACC)
    fbs = plan.info().FBS
    assert fbs == -(-(design.delaycount + batch - 1) // batch) * batch
    calls = -(-3 * fbs // (2 * batch))  # 1.5 ring depths of blocks
    n = per * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, dms[1], 0.0011, seed=150 + log2L)
    raw, x = np.ascontiguousarray(raw[:n]), x[:n]
    print(f"case sub-unit accumulate across the ring wrap: L=2^{log2L} dc={design.delaycount} ring={fbs} blocks "
          f"pushes={calls}", flush=True)
    outs = []
    for c in range(calls):
        plan.push(raw[c * per:(c + 1) * per])
        outs.append(plan.output_c64())
    plan.close()
    got = np.concatenate(outs, axis=1)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    wrap = fbs * design.M  # first sample whose block reuses a ring slot
    for d in range(len(dms)):
        e0, e1 = rel_l2(got[d, :wrap], ref[d, :wrap]), rel_l2(got[d, wrap:], ref[d, wrap:])
        print(f"  DM {dms[d]}: rel-L2 before the wrap {e0:.2e}, after {e1:.2e}", flush=True)
        assert e0 < TOL and e1 < TOL


@pytest.mark.parametrize("log2L", [17, 18])
def test_sub_accum_schedule_falls_back_for_fp16_output(log2L):
    # the sub-unit form has no FP16-pair epilogue (its per-block scale would span the Q sub-units): a plan
    # with FP16 output under the accumulate schedule runs the two-pass form, bit for bit as a plan
    # scheduled for it, FP16 pairs and block scales included
    design = _design(log2L)
    batch, calls = 2, 2
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, float(design.dms[1]), 0.0011,
                        seed=160 + log2L)
    raw = np.ascontiguousarray(raw[:n])
    res = []
    for order in (ACC, ms.ORDER_XCD_DM_MAJOR):
        print(f"case FP16 output L=2^{log2L} order={order}", flush=True)
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64 | ms.OUT_F16)
        plan.set_schedule(0, order)
        outs = []
        for c in range(calls):
            plan.push(raw[c * per:(c + 1) * per])
            outs.append((plan.output_c64(), *plan.output_f16()))
        plan.close()
        res.append(outs)
    for a, b in zip(res[0], res[1]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
