"""GPU parity of the DM-derivative pass (tangent push) against the FP64 oracle.

The tangent push dedisperses every block twice, with H and with dH/dDM, and folds
2 Re(conj(y) dy/dDM) with the primal fold's bins and skip rule, so the tangent fold's profile is
d profile / d DM. Bars: counts bit-exact (and equal to the primal fold's), primal profile within
rel-L2 1e-5, derivative profile within rel-L2 1e-4 (it is a sum of products that cancel to a few
percent of the primal power, so FP32 rounding is amplified ~10x relative to the series).
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import make_input, oracle_design, rel_l2

pytestmark = pytest.mark.gpu

BW, F0 = 16e6, 1392e6


@pytest.mark.parametrize("mode", ["persistent", "per_dm"])
def test_tangent_fold_matches_oracle(mode):
    dms = np.array([122.9, 123.456, 124.0], dtype=np.float32)
    batch, calls, nbins, period = 4, 3, 128, 0.0011
    design = ms.Design(BW, F0, dms)
    M = design.M
    per = batch * M
    n = per * calls
    raw, x = make_input(ms.IN_C64, 1 << int(np.ceil(np.log2(n))), BW, F0, 123.456, period, seed=17)
    raw, x = raw[:n], x[:n]
    print(f"case tangent mode={mode} M={M} dc={design.delaycount} batch={batch} calls={calls}", flush=True)
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_C64, out_flags=ms.OUT_NONE)
    if mode == "per_dm":
        plan.set_launch_mode(ms.LAUNCH_PER_DM)
    for i, dm in enumerate(dms):
        plan.set_dm_tangent(i, ms.design_transfer_function_ddm(BW, F0, float(dm), M))
    fold = ms.Fold(period, BW, per, nbins, num_series=len(dms))
    tfold = ms.Fold(period, BW, per, nbins, num_series=len(dms))
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
    delays, H = oracle_design(BW, F0, dms, M)
    dH = np.stack([orc.transfer_function_ddm(BW, F0, float(dm), M) for dm in dms])
    y = orc.msosm(x, M, delays, H)
    dy = orc.msosm(x, M, delays, dH)
    for d in range(len(dms)):
        p, dp, c = orc.profile_tangent(y[d], dy[d], nbins, 0.0, 1.0 / BW, period, call_len=per)
        assert np.array_equal(c, cnts[d])
        ep, edp = rel_l2(prof[d], p), rel_l2(dprof[d], dp)
        print(f"  DM {dms[d]}: profile rel-L2 {ep:.2e}, d profile/dDM rel-L2 {edp:.2e}", flush=True)
        assert ep < 1e-5
        assert edp < 1e-4


def test_autograd_dm_gradient_matches_oracle_and_differences():
    import torch
    from msosm_amd import autograd as ag

    dms = np.array([123.1, 123.456], dtype=np.float32)
    batch, calls, nbins, period = 4, 4, 128, 0.0011
    eng = ag.ProfileEngine(BW, F0, dms, batch, nbins, period, dm_max=float(dms.max()) + 0.01)
    M = eng.M
    n = batch * M * calls
    _, x = make_input(ms.IN_C64, 1 << int(np.ceil(np.log2(n))), BW, F0, 123.456, period, seed=23)
    x = x[:n]
    rec = torch.from_numpy(x.astype(np.complex64)).cuda()
    print(f"case autograd M={M} dc={eng.delaycount} n={n}", flush=True)
    noise_var = ag.profile_noise_var(rec, n, nbins)
    dm = torch.tensor(dms.astype(np.float64), requires_grad=True)
    loss = ag.snr_loss(ag.folded_profile(dm, eng, rec), noise_var)
    loss.sum().backward()
    g = dm.grad.numpy()
    # oracle: the same loss gradient from the FP64 profile and its DM derivative
    delays = np.stack([orc.segmentation(BW, F0, float(d), M)[0] for d in dms])
    H = np.stack([orc.transfer_function(BW, F0, float(d), M) for d in dms])
    dH = np.stack([orc.transfer_function_ddm(BW, F0, float(d), M) for d in dms])
    y, dy = orc.msosm(x, M, delays, H), orc.msosm(x, M, delays, dH)
    for i in range(len(dms)):
        p, dp, _ = orc.profile_tangent(y[i], dy[i], nbins, 0.0, 1.0 / BW, period, call_len=batch * M)
        ref = -2.0 * np.sum((p - p.mean()) * dp) / noise_var
        print(f"  DM {dms[i]}: grad {g[i]:.6e} oracle {ref:.6e}", flush=True)
        assert abs(g[i] - ref) <= 1e-4 * abs(ref) + 1e-9
    # central differences of the GPU loss. The loss varies on a DM scale of ~1e-3 here (a DM change of
    # 1e-3 leaves a residual chirp of ~2.4 rad at the band edge, which re-draws the noise in every
    # bin), so h = 1e-3 is far too coarse; h = 2^-14 keeps the truncation error near 1e-3 relative
    # and DM +- h exact in float32 (the DMs are float32 on the way to the design).
    h = 2.0 ** -14
    with torch.no_grad():
        lp = ag.snr_loss(ag.folded_profile(torch.tensor(dms + h, dtype=torch.float64), eng, rec), noise_var)
        lm = ag.snr_loss(ag.folded_profile(torch.tensor(dms - h, dtype=torch.float64), eng, rec), noise_var)
    fd = ((lp - lm) / (2 * h)).numpy()
    for i in range(len(dms)):
        print(f"  DM {dms[i]}: grad {g[i]:.6e} central difference {fd[i]:.6e}", flush=True)
        assert abs(g[i] - fd[i]) <= 2e-2 * abs(g[i]) + 1e-6


def test_loss_central_differences_at_configured_step():
    # Central differences of the loss at the configured step h = 1e-3, on the GPU and in the FP64 oracle
    # with the same designs at DM +- h (each designed anew at its DM, as the engine does: at DM 123.1 one
    # bin changes delay block within +-1e-3, in both). The two differences agree to the profiles'
    # rounding. The derivative pass's gradient is printed beside them: at this step the difference
    # also carries the loss's variation on DM scales below 1e-3 (the residual chirp re-draws the noise),
    # so it is a coarse check of the gradient, whose exact checks are the FP64 gradient and h = 2^-14.
    import torch
    from msosm_amd import autograd as ag

    dms = np.array([123.1, 123.456], dtype=np.float32)
    h = 1e-3
    batch, calls, nbins, period = 4, 4, 128, 0.0011
    eng = ag.ProfileEngine(BW, F0, dms, batch, nbins, period, dm_max=float(dms.max()) + 0.01)
    M = eng.M
    n = batch * M * calls
    _, x = make_input(ms.IN_C64, 1 << int(np.ceil(np.log2(n))), BW, F0, 123.456, period, seed=23)
    x = x[:n]
    rec = torch.from_numpy(x.astype(np.complex64)).cuda()
    noise_var = ag.profile_noise_var(rec, n, nbins)
    print(f"case central differences h={h} M={M} dc={eng.delaycount} n={n}", flush=True)

    def gpu_loss(d):
        with torch.no_grad():
            return ag.snr_loss(ag.folded_profile(torch.tensor(d, dtype=torch.float64), eng, rec), noise_var).numpy()

    def oracle_loss(d):
        out = []
        for v in d.astype(np.float32):
            dl, _ = orc.segmentation(BW, F0, float(v), M)
            H = orc.transfer_function(BW, F0, float(v), M)
            y = orc.msosm(x, M, dl[None], H[None])[0]
            s = np.zeros(nbins)
            c = np.zeros(nbins, dtype=np.uint64)
            t0 = 0.0
            for k in range(calls):
                orc.fold(np.abs(y[k * batch * M:(k + 1) * batch * M]) ** 2, nbins, t0, 1.0 / BW, period, s, c)
                t0 += float(batch * M) * (1.0 / BW)
            p = s / np.maximum(c, 1)
            out.append(-np.sum((p - p.mean()) ** 2) / noise_var)
        return np.array(out)

    dp, dm_ = (dms.astype(np.float64) + h).astype(np.float32), (dms.astype(np.float64) - h).astype(np.float32)
    fd_gpu = (gpu_loss(dp) - gpu_loss(dm_)) / (dp.astype(np.float64) - dm_.astype(np.float64))
    fd_orc = (oracle_loss(dp) - oracle_loss(dm_)) / (dp.astype(np.float64) - dm_.astype(np.float64))
    loss0 = oracle_loss(dms)
    dm_t = torch.tensor(dms.astype(np.float64), requires_grad=True)
    ag.snr_loss(ag.folded_profile(dm_t, eng, rec), noise_var).sum().backward()
    g = dm_t.grad.numpy()
    for i in range(len(dms)):
        changed = [int((orc.segmentation(BW, F0, float(v), M)[0] != orc.segmentation(BW, F0, float(dms[i]), M)[0]).sum())
                   for v in (dp[i], dm_[i])]
        print(f"  DM {dms[i]}: central difference h=1e-3 GPU {fd_gpu[i]:.6e} oracle {fd_orc[i]:.6e}; "
              f"gradient {g[i]:.6e}; bins changing block at +h/-h {changed}", flush=True)
        # each loss within ~4e-6 relative of the oracle's (profile rounding), divided by the 2h step
        tol = 4e-6 * abs(loss0[i]) / (2 * h) + 1e-3 * abs(fd_orc[i])
        assert abs(fd_gpu[i] - fd_orc[i]) <= tol


@pytest.mark.parametrize("seed", [31, 32])
def test_dm_search_reaches_fine_grid_optimum(seed):
    # DM refinement at reduced size: 2^22 samples of the refinement workload's pulsar (band 1392-1408 MHz,
    # DM 123.456, P = 33 ms, 3 % FWHM, amplitude 0.5, AWGN sigma 1), 512-bin folds, window +-0.05. The
    # search (a 1e-3 grid, then gradient refinement from its 4 deepest basins) ends at a loss no worse
    # than the best of the 1e-4 grid over the same window, within 1e-4 of that grid's best DM, with a
    # third of the grid's DM-record passes.
    import torch
    from msosm_amd import autograd as ag

    n, nbins, period, true, hw = 1 << 22, 512, 33e-3, 123.456, 0.05
    x = orc.pulsar(n, BW, F0, true, period, duty=0.03, amp=0.5, sigma=1.0, seed=seed).astype(np.complex64)
    rec = torch.from_numpy(x).cuda()
    nv = ag.profile_noise_var(rec, n, nbins)
    lo, hi = true - hw, true + hw
    print(f"case dm search seed={seed} n={n} window=[{lo}, {hi}]", flush=True)
    r = ag.search_dm(rec, n, BW, F0, lo, hi, nv, nbins, period, 32, 1e-3, polish=4, steps=20)
    grid = np.arange(lo, hi + 0.5e-4, 1e-4)
    gl = ag.grid_losses(rec, n, BW, F0, grid, nv, nbins, period, 32, float(np.float32(hi)))
    gb = int(np.argmin(gl))
    print(f"  search: DM {r['best_dm']:.6f} loss {r['best_loss']:.4f} in {r['passes']} passes; "
          f"1e-4 grid: DM {grid[gb]:.6f} loss {gl[gb]:.4f} in {grid.size} passes", flush=True)
    assert all(lo <= d <= hi for d in r["refined"])
    assert r["best_loss"] <= gl[gb]
    assert abs(r["best_dm"] - grid[gb]) < 1e-4
    assert r["passes"] * 2 < grid.size
