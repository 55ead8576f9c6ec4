"""The plain-OSM and precision-stress workloads at their full stated sizes, against the FP64 oracle.

* Plain single-DM OSM: one complex64 channel of 2^24 samples at 16 MS/s, band 1392-1408 MHz
  (fc 1400 MHz), DM 50, FFT length 2^18 (the OSM rule: 2 nextpow2(Nd), osm_prepare.cpp:17-30), a
  simulated pulsar (P = 33 ms, Gaussian pulse of 3 % FWHM, amplitude 0.5, AWGN sigma 1), folded to
  1024 bins. The whole record's series is compared with the oracle's, and the fold with the oracle's
  fold of the same series (counts bit-exact).
* Precision stress: 128 DMs 1500 (i + 1) / 128 at 325-375 MHz, 2^25 complex64 samples, FFT length 2^21
  (M = 2^20, delaycount 700: the record lies inside the dispersion window, zero history on both
  sides), chirp phase in FP64 and in the reference's FP32 form, output complex64 and FP16 pairs with a
  block scale. Three trials (i = 0, 63, 127) are compared with the oracle over the whole record; for all
  128 trials the FP16 output is compared with the complex64 output of the same launch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import device_view, oracle_design, rel_l2

pytestmark = pytest.mark.gpu


def _maxrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_plain_osm_full_record_against_oracle():
    fs, f0, dm, period, nbins, batch, n = 16e6, 1392e6, 50.0, 0.033, 1024, 4, 1 << 24
    M = orc.osm_min_order(orc.nd(fs, f0, dm))
    assert 2 * M == 1 << 18
    design = ms.Design(fs, f0, [dm], M=M)
    assert design.delaycount == 1 and not design.delays.any()  # one segment: MS-OSM is OSM
    per = batch * M
    print(f"case plain OSM full record: n={n} L={2 * M} pushes={n // per}", flush=True)
    x = orc.pulsar(n, fs, f0, dm, period, duty=0.03, amp=0.5, sigma=1.0, seed=20261203)
    raw = x.astype(np.complex64)
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64)
    fold = ms.Fold(period, fs, per, nbins, num_series=1)
    plan.attach_fold(fold)
    outs = []
    for c in range(n // per):
        plan.push(raw[c * per:(c + 1) * per])
        outs.append(plan.output_c64()[0])
    plan.synchronize()
    _, sums, cnts = fold.read()
    plan.close()
    fold.close()
    got = np.concatenate(outs)
    ref = orc.msosm(raw.astype(np.complex128), M, np.zeros((1, 2 * M), np.int32),
                    orc.transfer_function(fs, f0, float(np.float32(dm)), M)[None, :])[0]
    err, mx = rel_l2(got, ref), _maxrel(got, ref)
    rs, rc = np.zeros(nbins), np.zeros(nbins, dtype=np.uint64)
    t0, dt = 0.0, 1.0 / fs
    for c in range(n // per):
        orc.fold(np.abs(ref[c * per:(c + 1) * per]) ** 2, nbins, t0, dt, period, rs, rc)
        t0 += float(per) * dt
    srel = float(np.max(np.abs(sums[0] - rs) / rs))
    print(f"  series rel-L2 {err:.2e}, max |err|/max|y| {mx:.2e}; fold counts equal "
          f"{np.array_equal(cnts[0], rc)}, sum rel err {srel:.2e}", flush=True)
    assert err < 1e-5
    assert np.array_equal(cnts[0], rc)
    assert srel < 1e-5


def test_precision_stress_full_record_per_variant():
    bw, f0, n, batch = 50e6, 325e6, 1 << 25, 2
    dms = np.array([1500.0 * (i + 1) / 128 for i in range(128)], dtype=np.float32)
    pick = [0, 63, 127]
    rng = np.random.default_rng(2025)
    raw = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(0.5)).astype(np.complex64)
    res, f16_vs_c64 = {}, {}
    for phase, pname in ((ms.PHASE_FP64, "fp64-phase"), (ms.PHASE_FP32, "fp32-phase")):
        design = ms.Design(bw, f0, dms, phase_mode=phase)
        M = design.M
        assert M == 1 << 20 and design.delaycount == 700
        per = batch * M
        print(f"case precision full record {pname}: D=128 n={n} M={M} dc={design.delaycount} pushes={n // per}",
              flush=True)
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_C64 | ms.OUT_F16)
        c64_dev = device_view(plan.device_output_c64(), 128 * per * 2, "<f4").view(128, batch, M, 2)
        pf, ps = C.c_void_p(), C.c_void_p()
        assert ms.lib().msosm_plan_output_f16(plan.h, C.byref(pf), C.byref(ps)) == 0
        half_dev = device_view(pf.value, 128 * per * 2, "<f2").view(128, batch, M, 2)
        scale_dev = device_view(ps.value, 128 * batch, "<f4").view(128, batch, 1, 1)
        c64, f16 = [], []
        worst = 0.0
        for c in range(n // per):
            plan.push(raw[c * per:(c + 1) * per])
            plan.synchronize()
            # every trial: the FP16 pairs (times their block scale) against the complex64 series of the
            # same launch, on the device
            h = half_dev.float() * scale_dev
            num = (h - c64_dev).reshape(128, -1).norm(dim=1)
            den = c64_dev.reshape(128, -1).norm(dim=1)
            worst = max(worst, float((num / den).max()))
            to_c = lambda t: t[pick].reshape(len(pick), per, 2).cpu().numpy().view(np.complex64)[..., 0]  # noqa: E731
            c64.append(to_c(c64_dev))
            f16.append(to_c(h))
            del h
        plan.close()
        f16_vs_c64[phase] = worst
        res[phase] = (np.concatenate(c64, axis=1), np.concatenate(f16, axis=1))
        print(f"  {pname}: all 128 trials, FP16 pairs vs complex64 worst rel-L2 {worst:.2e}", flush=True)
        assert worst < 2.0 ** -11
    delays, H = oracle_design(bw, f0, dms[pick], 1 << 20)
    print("  FP64 oracle over the whole record (3 trials) ...", flush=True)
    ref = orc.msosm(raw.astype(np.complex128), 1 << 20, delays, H)
    for j, i in enumerate(pick):
        line = []
        for phase, pname in ((ms.PHASE_FP64, "fp64-phase"), (ms.PHASE_FP32, "fp32-phase")):
            for k, oname in ((0, "c64"), (1, "fp16")):
                y = res[phase][k][j]
                line.append(f"{pname}/{oname} rel-L2 {rel_l2(y, ref[j]):.2e} max {_maxrel(y, ref[j]):.2e}")
        print(f"  DM {dms[i]:.2f}: " + "; ".join(line), flush=True)
        assert rel_l2(res[ms.PHASE_FP64][0][j], ref[j]) < 1e-5
        assert rel_l2(res[ms.PHASE_FP64][1][j], ref[j]) < 2.0 ** -11
    worst32 = max(rel_l2(res[ms.PHASE_FP32][0][j], ref[j]) for j in range(len(pick)))
    worst64 = max(rel_l2(res[ms.PHASE_FP64][0][j], ref[j]) for j in range(len(pick)))
    assert worst32 > 10 * worst64


def test_dm_gradient_full_record():
    # The refinement workload at its stated size: 2^24 complex64 samples (band 1392-1408 MHz, pulsar DM
    # 123.456, P = 33 ms, 3 % FWHM, amplitude 0.5, AWGN sigma 1), 64 start DMs within +-0.5 of the truth,
    # loss -sum((p - mean p)^2) / noise variance over 512-bin profiles. The derivative pass's gradient
    # of every trial against central differences of the GPU loss (h = 2^-14: the loss varies on DM
    # scales near 1e-3 here, see test_gpu_gradient.py), and of two trials against the FP64 oracle's
    # gradient over the whole record.
    from msosm_amd import autograd as ag

    bw, f0, true, period, nbins, batch, n = 16e6, 1392e6, 123.456, 0.033, 512, 32, 1 << 24
    x = orc.pulsar(n, bw, f0, true, period, duty=0.03, amp=0.5, sigma=1.0, seed=2024)
    rec = torch.from_numpy(x.astype(np.complex64)).cuda()
    rng = np.random.default_rng(20261203)
    starts = (true + rng.uniform(-0.5, 0.5, 64)).astype(np.float32)
    eng = ag.ProfileEngine(bw, f0, starts, batch, nbins, period, dm_max=float(np.float32(true + 0.5)))
    M = eng.M
    print(f"case DM gradient full record: n={n} M={M} dc={eng.delaycount} trials=64", flush=True)
    nv = ag.profile_noise_var(rec, n, nbins)
    dm = torch.tensor(starts.astype(np.float64), requires_grad=True)
    ag.snr_loss(ag.folded_profile(dm, eng, rec), nv).sum().backward()
    g = dm.grad.numpy()
    h = 2.0 ** -14
    with torch.no_grad():
        lp = ag.snr_loss(ag.folded_profile(torch.tensor(starts + h, dtype=torch.float64), eng, rec), nv)
        lm = ag.snr_loss(ag.folded_profile(torch.tensor(starts - h, dtype=torch.float64), eng, rec), nv)
    fd = ((lp - lm) / (2 * h)).numpy()
    rel = np.abs(g - fd) / np.maximum(np.abs(g), 1e-12)
    print(f"  gradient vs central differences over 64 trials: median rel diff {np.median(rel):.2e}, "
          f"max {rel.max():.2e} (|g| median {np.median(np.abs(g)):.3e})", flush=True)
    assert np.median(rel) < 2e-2
    pick = [0, 37]
    xx = x.astype(np.complex64).astype(np.complex128)
    delays = np.stack([orc.segmentation(bw, f0, float(starts[i]), M)[0] for i in pick])
    H = np.stack([orc.transfer_function(bw, f0, float(starts[i]), M) for i in pick])
    dH = np.stack([orc.transfer_function_ddm(bw, f0, float(starts[i]), M) for i in pick])
    print("  FP64 oracle over the whole record (2 trials) ...", flush=True)
    y, dy = orc.msosm(xx, M, delays, H), orc.msosm(xx, M, delays, dH)
    for j, i in enumerate(pick):
        p, dp, _ = orc.profile_tangent(y[j], dy[j], nbins, 0.0, 1.0 / bw, period, call_len=batch * M)
        ref = -2.0 * np.sum((p - p.mean()) * dp) / nv
        print(f"  DM {starts[i]:.6f}: gradient {g[i]:.6e} oracle {ref:.6e} central difference {fd[i]:.6e}",
              flush=True)
        assert abs(g[i] - ref) <= 1e-4 * abs(ref) + 1e-9
