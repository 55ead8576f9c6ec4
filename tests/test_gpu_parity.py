"""GPU parity of the MS-OSM path against the FP64 oracle (small sizes the oracle finishes in seconds).

Bar: dedispersed series within rel-L2 1e-5 (FP32 arithmetic vs FP64); uint16 codes within one
code step; fold bin counts bit-exact; folded sums within 1e-5 relative.
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import check_u16_codes, gpu_series, make_input, oracle_design, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5

# cfg-1-like band (16 MHz at 1392-1408 MHz); M from the reference rule for the largest DM
BW, F0 = 16e6, 1392e6


def _run_case(fmt, dms, batch, calls, out=ms.OUT_C64, M=None, seed=3):
    design = ms.Design(BW, F0, dms, M=M)
    n = batch * design.M * calls
    n2 = 1 << int(np.ceil(np.log2(n)))
    raw, x = make_input(fmt, n2, BW, F0, dm=float(max(dms)), period=0.0011, seed=seed)
    raw, x = raw[:n], x[:n]
    print(f"case fmt={fmt} dms={list(dms)} M={design.M} dc={design.delaycount} batch={batch} calls={calls}",
          flush=True)
    got = gpu_series(design, raw, fmt, batch, out)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    assert np.array_equal(delays, design.delays)
    ref = orc.msosm(x, design.M, delays, H)
    return design, got, ref


@pytest.mark.parametrize("fmt", [ms.IN_C64, ms.IN_I8, ms.IN_U16])
def test_series_c64_matches_oracle(fmt):
    design, got, ref = _run_case(fmt, [0.0, 20.0, 50.0], batch=4, calls=3)
    for d in range(got.shape[0]):
        err = rel_l2(got[d], ref[d])
        print(f"  DM {design.dms[d]}: rel-L2 {err:.3e}", flush=True)
        assert err < TOL


def test_dm0_is_identity_delay_of_one_block():
    # DM 0: H == 1 and no delays, so out[t] = x[t] exactly (up to FFT rounding)
    design, got, ref = _run_case(ms.IN_C64, [0.0], batch=2, calls=2, M=4096)
    assert rel_l2(got[0], ref[0]) < TOL


def test_uint16_output_codes():
    design, got, ref = _run_case(ms.IN_U16, [10.0, 50.0], batch=4, calls=2, out=ms.OUT_U16)
    # codes equal the oracle's encoding of the FP64 series except at rounding ties
    frac, worst = check_u16_codes(got, ref)
    print(f"  max code diff {worst}, frac differing (all at ties) {frac:.2e}", flush=True)


def test_fused_fold_matches_oracle():
    # DM 0 is left out: its output equals the input up to rounding, so integer-zero input samples
    # sit at the I == 0 skip threshold where FP32 and FP64 legitimately disagree
    dms = [5.0, 25.0, 50.0]
    batch, calls, nbins, period = 4, 3, 256, 0.0011
    design = ms.Design(BW, F0, dms)
    n = batch * design.M * calls
    raw, x = make_input(ms.IN_I8, 1 << int(np.ceil(np.log2(n))), BW, F0, 50.0, period, seed=11)
    raw, x = raw[:n], x[:n]
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_NONE)
    fold = ms.Fold(period, BW, batch * design.M, nbins, num_series=len(dms))
    plan.attach_fold(fold)
    per = batch * design.M
    for c in range(calls):
        plan.push(raw[c * per:(c + 1) * per])
    plan.synchronize()
    prof, sums, cnts = fold.read()
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    dt = 1.0 / BW
    for d in range(len(dms)):
        rs = np.zeros(nbins)
        rc = np.zeros(nbins, dtype=np.uint64)
        t0 = 0.0
        for c in range(calls):
            seg = np.abs(ref[d, c * per:(c + 1) * per]) ** 2
            orc.fold(seg, nbins, t0, dt, period, rs, rc)
            t0 += float(per) * dt  # t_start += size * dt (fold_gpu.cpp:53)
        print(f"  DM {dms[d]}: counts equal {np.array_equal(cnts[d], rc)}, sum rel err "
              f"{np.max(np.abs(sums[d] - rs) / rs):.2e}", flush=True)
        assert np.array_equal(cnts[d], rc)
        assert np.max(np.abs(sums[d] - rs) / rs) < TOL
        np.testing.assert_allclose(prof[d], rs / rc, rtol=TOL)
    plan.close()
    fold.close()


@pytest.mark.parametrize("log2L", list(range(12, 23)))
def test_every_fft_length(log2L):
    # every compiled (N1, N2) shape: one DM whose smear spans a few blocks at this M
    M = 1 << (log2L - 1)
    dm = 30.0 * M / 8192
    design = ms.Design(BW, F0, [dm], M=M)
    calls = 3
    n = design.M * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, dm, 0.0011, seed=log2L)
    raw, x = raw[:n], x[:n]
    print(f"L=2^{log2L} dc={design.delaycount}", flush=True)
    got = gpu_series(design, raw, ms.IN_C64, 1)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    err = rel_l2(got[0], ref[0])
    print(f"  rel-L2 {err:.3e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("batch", [1, 3])
def test_ring_wraps_many_times(batch):
    # FBS = roundup(dc + batch - 1, batch) slots; 24 blocks wrap the ring several times
    design = ms.Design(BW, F0, [6.0, 10.0], M=2048)
    calls = 24 // batch
    n = batch * design.M * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 10.0, 0.0011, seed=batch)
    raw, x = raw[:n], x[:n]
    got = gpu_series(design, raw, ms.IN_C64, batch)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    for d in range(2):
        assert rel_l2(got[d], ref[d]) < TOL


def test_fp32_phase_variant_runs_and_is_less_accurate():
    # same data through the FP64-phase and the legacy FP32-phase design: the FP32 phase is off
    dms = [50.0]
    d64 = ms.Design(BW, F0, dms)
    d32 = ms.Design(BW, F0, dms, phase_mode=ms.PHASE_FP32)
    n = 2 * d64.M * 2
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 50.0, 0.0011, seed=9)
    raw, x = raw[:n], x[:n]
    delays, H = oracle_design(BW, F0, d64.dms, d64.M)
    ref = orc.msosm(x, d64.M, delays, H)
    e64 = rel_l2(gpu_series(d64, raw, ms.IN_C64, 2)[0], ref[0])
    e32 = rel_l2(gpu_series(d32, raw, ms.IN_C64, 2)[0], ref[0])
    print(f"  FP64 phase rel-L2 {e64:.2e}, FP32 phase rel-L2 {e32:.2e}", flush=True)
    assert e64 < TOL and e32 > 10 * e64


def test_fp16_pair_output_with_block_scales():
    """FP16-pair output: each (DM, block) carries a power-of-two scale at or below the rms of the
    block's length-2M inverse transform; the pairs are y / scale rounded to nearest even."""
    dms = [0.0, 20.0, 50.0]
    batch, calls = 4, 3
    design = ms.Design(BW, F0, dms)
    per = batch * design.M
    n = per * calls
    raw, x = make_input(ms.IN_I8, 1 << int(np.ceil(np.log2(n))), BW, F0, 50.0, 0.0011, seed=5)
    raw, x = raw[:n], x[:n]
    got, got_c64, scl = [], [], []
    for mode in (ms.LAUNCH_PERSISTENT, ms.LAUNCH_PER_DM):
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_C64 | ms.OUT_F16)
        plan.set_launch_mode(mode)
        v, c, s = [], [], []
        for k in range(calls):
            plan.push(raw[k * per:(k + 1) * per])
            vals, scales = plan.output_f16()
            v.append(vals)
            c.append(plan.output_c64())
            s.append(scales)
        plan.close()
        got.append(np.concatenate(v, axis=1))
        got_c64.append(np.concatenate(c, axis=1))
        scl.append(np.concatenate(s, axis=1))
    # both launch forms write identical pairs and scales
    assert np.array_equal(got[0], got[1]) and np.array_equal(scl[0], scl[1])
    vals, c64, scales = got[0], got_c64[0], scl[0]
    mant, _ = np.frexp(scales)
    assert np.all(mant == 0.5), "scales are powers of two"
    # the codec is exact relative to the same kernel's complex64 series
    s_full = np.repeat(scales, design.M, axis=1)
    enc = lambda r: np.clip(r / s_full, -65504, 65504).astype(np.float16).astype(np.float32) * s_full
    assert np.array_equal(vals.real, enc(c64.real)) and np.array_equal(vals.imag, enc(c64.imag))
    # the scale follows the block's power
    rms = np.sqrt(np.mean(np.abs(c64.reshape(len(dms), -1, design.M)) ** 2, axis=2))
    ratio = rms / scales
    print(f"  rms/scale in [{ratio.min():.2f}, {ratio.max():.2f}]", flush=True)
    assert np.all(ratio > 0.25) and np.all(ratio < 4.0)
    # against the FP64 oracle: FP16 keeps 11 significant bits, unit roundoff 2^-11 per component
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    for d in range(len(dms)):
        err = rel_l2(vals[d], ref[d])
        print(f"  DM {dms[d]}: FP16 rel-L2 {err:.3e}", flush=True)
        assert err < 2.0 ** -11


def test_schedules_are_bitwise_identical_and_autotune_resets():
    # the tile schedule changes which workgroup runs which unit, never the arithmetic
    import torch

    dms = [0.0, 20.0, 50.0]
    batch, calls = 4, 3
    design = ms.Design(BW, F0, dms)
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 50.0, 0.0011, seed=41)
    raw = np.ascontiguousarray(raw[:n])
    dev = torch.from_numpy(raw.view(np.float32)).cuda()
    outs = []
    for wg, order in [(0, ms.ORDER_XCD_DM_MAJOR), (0, ms.ORDER_LINEAR), (0, ms.ORDER_XCD_BLOCK_MAJOR),
                      (0, ms.ORDER_XCD_TILE4), (0, ms.ORDER_XCD_TILE8), (0, ms.ORDER_XCD_TILE16),
                      (0, ms.ORDER_XCD_TILE2), (16, 0), (64, ms.ORDER_XCD_TILE8), (128, ms.ORDER_XCD_TILE16),
                      (0, ms.ORDER_LINEAR | ms.SCHEDULE_SPLIT), (8, ms.ORDER_XCD_DM_MAJOR | ms.SCHEDULE_SPLIT)]:
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
        plan.set_schedule(wg, order)
        o = []
        for c in range(calls):
            plan.push(dev.data_ptr() + c * per * 8)
            o.append(plan.output_c64())
        outs.append(np.concatenate(o, axis=1))
        plan.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
    w, order, t = plan.autotune(dev.data_ptr(), calls)
    print(f"autotune: workgroups {w} order {order} {t:.3f} ms", flush=True)
    assert plan.schedule() == (w, order) and t > 0
    o = []
    for c in range(calls):  # the tuning pushes were undone: the stream starts from zero history again
        plan.push(dev.data_ptr() + c * per * 8)
        o.append(plan.output_c64())
    plan.close()
    # the same bits as a fresh plan on the chosen schedule (the accumulate form is another
    # factorisation of the transform: equal to the others to rounding, to itself bit for bit)
    ref = outs[0]
    if order & ms.SCHEDULE_ACCUM:
        fresh = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
        fresh.set_schedule(w, order)
        r = []
        for c in range(calls):
            fresh.push(dev.data_ptr() + c * per * 8)
            r.append(fresh.output_c64())
        fresh.close()
        ref = np.concatenate(r, axis=1)
    assert np.array_equal(ref, np.concatenate(o, axis=1))


@pytest.mark.parametrize("band", ["L2^14", "L2^16"])
def test_role_split_matches_sequential(band):
    # the role-split launch (one half of each workgroup on pass A of a unit, the other on pass B of the
    # previous unit, alternating slot layouts) against the sequential form: identical series bits and
    # fold counts, with workgroups running up to 12 units each (every layout change, and every unit
    # of a push chained through one workgroup's look-ahead)
    import torch

    bw, f0, dms = (BW, F0, [5.0, 25.0, 50.0]) if band == "L2^14" else (64e6, 1218e6, [10.0, 50.0, 100.0])
    batch, calls, nbins, period = 4, 3, 256, 0.0011
    design = ms.Design(bw, f0, dms)
    assert 2 * design.M == (1 << 14 if band == "L2^14" else 1 << 16)
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), bw, f0, float(max(dms)), period, seed=43)
    dev = torch.from_numpy(np.ascontiguousarray(raw[:n]).view(np.float32)).cuda()
    res = []
    for wg, order in [(8, ms.ORDER_LINEAR), (8, ms.ORDER_LINEAR | ms.SCHEDULE_SPLIT),
                      (1, ms.ORDER_LINEAR | ms.SCHEDULE_SPLIT), (16, ms.ORDER_XCD_DM_MAJOR | ms.SCHEDULE_SPLIT),
                      (0, ms.ORDER_XCD_TILE4 | ms.SCHEDULE_SPLIT)]:
        print(f"case {band} workgroups={wg} order={order}", flush=True)
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
        fold = ms.Fold(period, bw, per, nbins, num_series=len(dms))
        plan.attach_fold(fold)
        plan.set_schedule(wg, order)
        o = []
        for c in range(calls):
            plan.push(dev.data_ptr() + c * per * 8)
            o.append(plan.output_c64())
        _, sums, cnts = fold.read()
        res.append((np.concatenate(o, axis=1), sums, cnts))
        plan.close()
        fold.close()
    ref = res[0]
    for series, sums, cnts in res[1:]:
        assert np.array_equal(series, ref[0])
        assert np.array_equal(cnts, ref[2])
        np.testing.assert_allclose(sums, ref[1], rtol=1e-12)  # FP64 sums: atomic order only


ORDER_NAMES = {ms.ORDER_XCD_DM_MAJOR: "xcd_dm_major", ms.ORDER_LINEAR: "linear",
               ms.ORDER_XCD_BLOCK_MAJOR: "xcd_block_major", ms.ORDER_XCD_TILE4: "xcd_tile4",
               ms.ORDER_XCD_TILE8: "xcd_tile8", ms.ORDER_XCD_TILE16: "xcd_tile16", ms.ORDER_XCD_TILE2: "xcd_tile2"}


@pytest.mark.parametrize("band", ["L2^14", "L2^16"])
def test_role_split_every_autotune_candidate(band):
    # every schedule the autotuner can pick at these lengths (each unit order, on all CUs and on half of
    # them, split or not) against the sequential linear launch: identical series bits and fold counts.
    # 40 trials x 8 blocks = 320 units per push, more than one per workgroup, so the tile orders' edge
    # positions (trials beyond the grid) and the role split's look-ahead across units both occur.
    import torch

    bw, f0 = (BW, F0) if band == "L2^14" else (64e6, 1218e6)
    dms = np.linspace(0.0, 50.0 if band == "L2^14" else 100.0, 40).astype(np.float32)
    batch, calls, nbins, period = 8, 2, 512, 0.0011
    design = ms.Design(bw, f0, dms)
    assert 2 * design.M == (1 << 14 if band == "L2^14" else 1 << 16)
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), bw, f0, float(dms[-1]), period, seed=47)
    dev = torch.from_numpy(np.ascontiguousarray(raw[:n]).view(np.float32)).cuda()

    def run(wg, order):
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
        fold = ms.Fold(period, bw, per, nbins, num_series=len(dms))
        plan.attach_fold(fold)
        plan.set_schedule(wg, order)
        o = []
        for c in range(calls):
            plan.push(dev.data_ptr() + c * per * 8)
            o.append(plan.output_c64())
        _, sums, cnts = fold.read()
        G = plan.info().workgroups
        plan.close()
        fold.close()
        return np.concatenate(o, axis=1), sums, cnts, G

    ref = run(0, ms.ORDER_LINEAR)
    G = ref[3]
    half = max(8, (G // 2 // 8) * 8)
    for wg in (G, half):
        for order, name in ORDER_NAMES.items():
            for split in (0, ms.SCHEDULE_SPLIT):
                print(f"case {band} workgroups={wg} order={name}{'+split' if split else ''}", flush=True)
                series, sums, cnts, _ = run(wg, order | split)
                assert np.array_equal(series, ref[0]), f"{name} split={bool(split)} wg={wg}: series differ"
                assert np.array_equal(cnts, ref[2])
                np.testing.assert_allclose(sums, ref[1], rtol=1e-12)


@pytest.mark.parametrize("fmt", [ms.IN_C64, ms.IN_I8, ms.IN_U16])
@pytest.mark.parametrize("log2L", [15, 16, 18])
def test_split_forward_transform_matches_one_workgroup_form(log2L, fmt):
    # a push of fewer blocks than the GPU holds workgroups splits each block's forward transform over
    # several workgroups (two launches); with as many workgroups as blocks it runs one workgroup per
    # block. The tiles and their arithmetic are the same, so the series are the same bits.
    design = ms.Design(BW, F0, [20.0, 60.0], M=1 << (log2L - 1))
    batch, calls = 3, 2
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(fmt, 1 << (n - 1).bit_length(), BW, F0, 60.0, 0.0011, seed=60 + log2L)
    raw = np.ascontiguousarray(raw[:n])
    res = []
    for wg in (batch, 0):
        print(f"case forward transform L=2^{log2L} fmt={fmt} batch={batch} workgroups={wg or 'all CUs'}", flush=True)
        plan = ms.Plan.from_design(design, batch, in_format=fmt, out_flags=ms.OUT_C64)
        plan.set_schedule(wg, ms.ORDER_XCD_DM_MAJOR)
        outs = []
        for c in range(calls):
            plan.push(raw[c * per:(c + 1) * per])
            outs.append(plan.output_c64())
        plan.close()
        res.append(np.concatenate(outs, axis=1))
    assert np.array_equal(res[0], res[1])
