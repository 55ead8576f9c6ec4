"""Full-size GPU checks at the benchmark geometries, through size-independent properties.

The FP64 oracle cannot run these sizes in seconds, so the whole records are checked by properties
that hold exactly at any size, plus a cross-check of a few trials against the FP32 CPU port (itself
pinned against the FP64 oracle in test_cpu_port.py):

* fold counts equal the host's count of samples per bin under the reference's bin expression
  (gpu_kernel.cu:477-481 with t_start advanced by size*dt per call, fold_gpu.cpp:53), for every trial
  whose series has no exact zero;
* scaling the input by 2 scales every series by 2 exactly (FP32 arithmetic commutes with power-of-two
  scaling away from overflow and underflow), so no count changes and every folded sum scales by 4 up
  to the order of its FP64 atomic additions (reassociation, ~1e-13 relative; any FP32-level
  difference would be ~1e-7);
* the folded sums of sampled trials agree with the CPU port's to FP32 rounding.
"""
import numpy as np
import pytest
import torch

import msosm_amd as ms
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def host_counts(n_calls, per_call, nbins, fs, period):
    """Samples per fold bin over n_calls calls of per_call samples, under the reference expression."""
    dt = 1.0 / fs
    i = np.arange(per_call, dtype=np.float64)
    counts = np.zeros(nbins, dtype=np.uint64)
    t0 = 0.0
    for _ in range(n_calls):
        t = t0 + i * dt  # two roundings, no contraction, as the kernels
        b = np.floor(np.fmod(t / period, 1.0) * nbins).astype(np.int64)
        np.minimum(b, nbins - 1, out=b)
        counts += np.bincount(b, minlength=nbins).astype(np.uint64)
        t0 += float(per_call) * dt
    return counts


def scaled_rel(a, b):
    """Largest |a - b| / |b| over bins (empty bins compare as absolute differences)."""
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def run_grid(design, q_dev, batch, nbins, fs, period, scale):
    """The whole int8 record (device tensor, already scaled) through a fused-fold plan."""
    per = batch * design.M
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_NONE, device=0)
    fold = ms.Fold(period, fs, per, nbins, num_series=len(design.dms), device=0)
    plan.attach_fold(fold)
    src = (q_dev * scale) if scale != 1 else q_dev
    torch.cuda.synchronize()
    calls = src.numel() // (2 * per)
    for c in range(calls):
        plan.push(src.data_ptr() + c * 2 * per)
    plan.synchronize()
    _, sums, cnts = fold.read()
    plan.close()
    fold.close()
    return sums, cnts, calls


def test_dm_grid_int8_full_record():
    # 512 trials over 0-100 pc/cm^3, band 1218-1282 MHz, 2^27 int8 samples at 64 MS/s, L = 2^16,
    # 32 blocks per push, every trial folded to 1024 bins at P = 5 ms; I/Q uniform in [-63, 63] so
    # that the doubled record is still int8
    fs, f0, nbins, period, batch, n = 64e6, 1218e6, 1024, 5.0e-3, 32, 1 << 27
    design = ms.Design(fs, f0, np.linspace(0.0, 100.0, 512))
    assert design.M == 1 << 15
    rng = np.random.default_rng(2026)
    q = rng.integers(-63, 64, size=2 * n, dtype=np.int8)
    q_dev = torch.from_numpy(q).to("cuda:0")
    print(f"dm grid: M={design.M} dc={design.delaycount} pushes={n // (batch * design.M)}", flush=True)
    sums1, cnts1, calls = run_grid(design, q_dev, batch, nbins, fs, period, 1)
    sums2, cnts2, _ = run_grid(design, q_dev, batch, nbins, fs, period, 2)
    # exact scaling: |2y|^2 = 4|y|^2 in FP32; the FP64 sums differ only by atomic-order reassociation
    assert np.array_equal(cnts1, cnts2)
    assert scaled_rel(sums2, 4.0 * sums1) < 1e-11
    # counts: every sample of every trial but DM 0 (whose series is the input delayed by one block,
    # so exact-zero input pairs sit at the I == 0 skip threshold)
    ref_counts = host_counts(calls, batch * design.M, nbins, fs, period)
    assert int(ref_counts.sum()) == n
    bad = [d for d in range(1, 512) if not np.array_equal(cnts1[d], ref_counts)]
    assert not bad, f"fold counts differ for trials {bad[:8]}"
    # sampled trials against the FP32 CPU port over the whole record
    pick = np.linspace(1, 511, 8).astype(int)
    x = orc.int8_to_complex(q).astype(np.complex64)
    cs, cc, _ = orc.cpu_msosm_fold(x, design.M, design.delays[pick], design.H[pick], nbins, 0.0, 1.0 / fs, period,
                                   call_size=batch * design.M)
    for j, d in enumerate(pick):
        assert np.array_equal(cnts1[d], cc[j])
        rel = scaled_rel(sums1[d], cs[j])
        print(f"  DM {design.dms[d]:.3f}: max bin rel diff vs CPU port {rel:.2e}", flush=True)
        assert rel < 1e-5


def test_dm_grid_role_split_full_record_matches_sequential():
    # The launch the benchmark times, at its shape: 512 trials over 0-100 pc/cm^3 at L = 2^16, 32 blocks
    # per push over the whole 2^27-sample int8 record (I/Q uniform in [-127, 127]), on every CU, in the
    # role-split form with 16-DM XCD tiles. Against the sequential launch of the same tile order, every
    # push's complex64 series (all 512 x 2^20 values, compared on the device) is bit-identical, and the
    # fused 1024-bin folds have identical counts and sums equal up to the order of their FP64 atomics.
    from tests.util import device_view

    fs, f0, nbins, period, batch, n = 64e6, 1218e6, 1024, 5.0e-3, 32, 1 << 27
    design = ms.Design(fs, f0, np.linspace(0.0, 100.0, 512).astype(np.float32))
    assert design.M == 1 << 15 and design.delaycount == 54
    per = batch * design.M
    g = torch.Generator(device="cuda:0").manual_seed(1234)
    q_dev = torch.randint(-127, 128, (2 * n,), dtype=torch.int8, device="cuda:0", generator=g)
    plans, folds, views = [], [], []
    for order in (ms.ORDER_XCD_TILE16 | ms.SCHEDULE_SPLIT, ms.ORDER_XCD_TILE16):
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_C64, device=0)
        fold = ms.Fold(period, fs, per, nbins, num_series=512, device=0)
        plan.attach_fold(fold)
        plan.set_schedule(0, order)
        plans.append(plan)
        folds.append(fold)
        views.append(device_view(plan.device_output_c64(), 512 * per * 2, "<u4"))
    wg, order = plans[0].schedule()
    assert wg == plans[0].info().workgroups and order & ms.SCHEDULE_SPLIT
    print(f"role split full record: M={design.M} dc={design.delaycount} workgroups={wg} pushes={n // per}",
          flush=True)
    torch.cuda.synchronize()
    for c in range(n // per):
        for plan in plans:
            plan.push(q_dev.data_ptr() + c * 2 * per)
        for plan in plans:
            plan.synchronize()
        if not torch.equal(views[0], views[1]):
            diff = int((views[0] != views[1]).sum())
            raise AssertionError(f"push {c}: {diff} series words differ between role split and sequential")
    _, s0, c0 = folds[0].read()
    _, s1, c1 = folds[1].read()
    assert np.array_equal(c0, c1)
    np.testing.assert_allclose(s0, s1, rtol=1e-12)
    for p_, f_ in zip(plans, folds):
        p_.close()
        f_.close()


def test_uneven_grid_full_record():
    # 1024 trials over 0-1000 pc/cm^3, band 784-816 MHz, 2^26 complex64 samples at 32 MS/s, each trial
    # at its own FFT length (2^14..2^18 here), one launch over all lengths per push of 4 x 2^17
    # samples, every trial folded to 1024 bins at P = 33 ms
    fs, f0, nbins, period, n = 32e6, 784e6, 1024, 33.0e-3, 1 << 26
    design = ms.UnevenDesign(fs, f0, np.linspace(0.0, 1000.0, 1024))
    g = torch.Generator(device="cuda:0").manual_seed(2027)
    x_dev = torch.randn(2 * n, generator=g, device="cuda:0", dtype=torch.float32)
    sums, cnts = [], []
    for scale in (1, 2):
        multi = ms.Multi(design, push_blocks=4, out_flags=ms.OUT_NONE, device=0)
        multi.fold(period, fs, nbins)
        src = x_dev * scale if scale != 1 else x_dev
        torch.cuda.synchronize()
        per = multi.push_samples
        calls = n // per
        print(f"uneven: classes={multi.num_classes} push={per} pushes={calls} scale={scale}", flush=True)
        for c in range(calls):
            multi.push(src.data_ptr() + c * per * 8)
        multi.synchronize()
        _, s, k = multi.fold_read()
        sums.append(s)
        cnts.append(k)
        multi.close()
    assert np.array_equal(cnts[0], cnts[1])
    assert scaled_rel(sums[1], 4.0 * sums[0]) < 1e-11
    ref_counts = host_counts(calls, per, nbins, fs, period)
    bad = [d for d in range(1, 1024) if not np.array_equal(cnts[0][d], ref_counts)]
    assert not bad, f"fold counts differ for trials {bad[:8]}"
    # one trial of each length class against the FP32 CPU port
    x = x_dev.cpu().numpy().view(np.complex64)
    for M in np.unique(design.M):
        d = int(np.nonzero(design.M == M)[0][-1])
        cs, cc, _ = orc.cpu_msosm_fold(x, int(M), design.delays[d][None], design.H[d][None], nbins, 0.0, 1.0 / fs,
                                       period, call_size=per)
        assert np.array_equal(cnts[0][d], cc[0])
        rel = scaled_rel(sums[0][d], cs[0])
        print(f"  DM {design.dms[d]:.1f} (L = {2 * M}): max bin rel diff vs CPU port {rel:.2e}", flush=True)
        assert rel < 1e-5


def test_dm_grid_accumulate_full_record_matches_other_orders_and_two_pass():
    # The launch the benchmark times (the accumulate form in 16-DM XCD tiles on every CU), at its shape:
    # 512 trials over 0-100 pc/cm^3 at L = 2^16, 32 blocks per push over the whole 2^27-sample int8 record.
    # Against the accumulate form in the linear unit order on half the CUs, every push's complex64 series
    # (all 512 x 2^20 values, compared on the device) is bit-identical and the fused folds have identical
    # counts and sums equal up to the order of their FP64 atomics. Against the sequential two-pass launch
    # (another factorisation of the same transform) each push's series agrees to FP32 rounding and the
    # fold counts are identical for every trial but DM 0 (an identity delay whose exact-zero input pairs
    # sit at the I == 0 skip threshold).
    from tests.util import device_view

    fs, f0, nbins, period, batch, n = 64e6, 1218e6, 1024, 5.0e-3, 32, 1 << 27
    design = ms.Design(fs, f0, np.linspace(0.0, 100.0, 512).astype(np.float32))
    assert design.M == 1 << 15 and design.delaycount == 54
    per = batch * design.M
    g = torch.Generator(device="cuda:0").manual_seed(4321)
    q_dev = torch.randint(-127, 128, (2 * n,), dtype=torch.int8, device="cuda:0", generator=g)
    plans, folds, words, values = [], [], [], []
    schedules = [(0, ms.ORDER_XCD_TILE16 | ms.SCHEDULE_ACCUM), (128, ms.ORDER_LINEAR | ms.SCHEDULE_ACCUM),
                 (0, ms.ORDER_XCD_TILE16)]
    for wg, order in schedules:
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_C64, device=0)
        fold = ms.Fold(period, fs, per, nbins, num_series=512, device=0)
        plan.attach_fold(fold)
        plan.set_schedule(wg, order)
        plans.append(plan)
        folds.append(fold)
        words.append(device_view(plan.device_output_c64(), 512 * per * 2, "<u4"))
        values.append(device_view(plan.device_output_c64(), 512 * per * 2, "<f4"))
    wg, order = plans[0].schedule()
    assert wg == plans[0].info().workgroups and order & ms.SCHEDULE_ACCUM
    print(f"accumulate full record: M={design.M} dc={design.delaycount} workgroups={wg} pushes={n // per}",
          flush=True)
    torch.cuda.synchronize()
    worst = 0.0
    for c in range(n // per):
        for plan in plans:
            plan.push(q_dev.data_ptr() + c * 2 * per)
        for plan in plans:
            plan.synchronize()
        if not torch.equal(words[0], words[1]):
            diff = int((words[0] != words[1]).sum())
            raise AssertionError(f"push {c}: {diff} series words differ between unit orders of the accumulate form")
        a, b = values[0].view(512, -1), values[2].view(512, -1)
        rel = float(((a - b).norm(dim=1) / b.norm(dim=1)).max())
        worst = max(worst, rel)
        assert rel < 2e-6, f"push {c}: accumulate vs two-pass rel-L2 {rel:.3e}"
    print(f"  accumulate vs two-pass: worst per-trial rel-L2 over all pushes {worst:.2e}", flush=True)
    _, s0, c0 = folds[0].read()
    _, s1, c1 = folds[1].read()
    _, s2, c2 = folds[2].read()
    assert np.array_equal(c0, c1)
    np.testing.assert_allclose(s0, s1, rtol=1e-12)
    assert np.array_equal(c0[1:], c2[1:])
    assert scaled_rel(s0[1:], s2[1:]) < 1e-5
    for p_, f_ in zip(plans, folds):
        p_.close()
        f_.close()
