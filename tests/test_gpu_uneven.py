"""Uneven DM grids (each DM at its own MS-OSM length) and the one-launch-per-DM form, against the
FP64 oracle evaluated per DM at that DM's length."""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import make_input, rel_l2

pytestmark = pytest.mark.gpu

BW, F0 = 16e6, 1392e6
TOL = 1e-5


def _oracle(design, x, i):
    M = int(design.M[i])
    dm = float(design.dms[i])
    d, _ = orc.segmentation(BW, F0, dm, M)
    H = orc.transfer_function(BW, F0, dm, M)
    return orc.msosm(x, M, d[None], H[None])[0]


def test_uneven_lengths_match_oracle_per_dm():
    design = ms.UnevenDesign(BW, F0, [10.0, 200.0, 800.0, 0.0, 450.0])
    assert len(set(design.M.tolist())) >= 3  # several length classes
    multi = ms.Multi(design, push_blocks=2)
    calls = 3
    n = multi.push_samples * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 450.0, 0.0011, seed=31)
    raw, x = raw[:n], x[:n]
    outs = [[] for _ in range(multi.D)]
    for c in range(calls):
        multi.push(raw[c * multi.push_samples:(c + 1) * multi.push_samples])
        for i in range(multi.D):
            outs[i].append(multi.output(i))
    print(f"classes {multi.num_classes}, M per DM {design.M.tolist()}", flush=True)
    for i in range(multi.D):
        got = np.concatenate(outs[i])
        err = rel_l2(got, _oracle(design, x, i))
        print(f"  DM {design.dms[i]} (M={design.M[i]}): rel-L2 {err:.2e}", flush=True)
        assert err < TOL
    multi.close()


def test_uneven_fold_counts_match_oracle():
    design = ms.UnevenDesign(BW, F0, [30.0, 300.0, 700.0])
    multi = ms.Multi(design, push_blocks=2, out_flags=ms.OUT_NONE)
    nbins, period = 128, 0.0011
    multi.fold(period, BW, nbins)
    calls = 2
    n = multi.push_samples * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 300.0, period, seed=32)
    raw, x = raw[:n], x[:n]
    for c in range(calls):
        multi.push(raw[c * multi.push_samples:(c + 1) * multi.push_samples])
    multi.synchronize()
    prof, sums, cnts = multi.fold_read()
    dt = 1.0 / BW
    for i in range(multi.D):
        ref = _oracle(design, x, i)
        rs, rc = np.zeros(nbins), np.zeros(nbins, dtype=np.uint64)
        t0 = 0.0
        for c in range(calls):
            seg = ref[c * multi.push_samples:(c + 1) * multi.push_samples]
            orc.fold(np.abs(seg) ** 2, nbins, t0, dt, period, rs, rc)
            t0 += float(multi.push_samples) * dt
        assert np.array_equal(cnts[i], rc)
        assert np.max(np.abs(sums[i] - rs) / rs) < TOL
    multi.close()


def test_per_dm_launch_form_is_bitwise_identical():
    # the comparison form (one launch per DM) runs the same arithmetic: outputs equal bit for bit
    design = ms.Design(BW, F0, [5.0, 25.0, 50.0])
    batch = 4
    n = batch * design.M * 2
    raw, _ = make_input(ms.IN_I8, 1 << (n - 1).bit_length(), BW, F0, 50.0, 0.0011, seed=33)
    raw = raw[:n]
    res = []
    for mode in (ms.LAUNCH_PERSISTENT, ms.LAUNCH_PER_DM):
        plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_C64)
        plan.set_launch_mode(mode)
        o = []
        for c in range(2):
            plan.push(raw[c * batch * design.M:(c + 1) * batch * design.M])
            o.append(plan.output_c64())
        res.append(np.concatenate(o, axis=1))
        plan.close()
    assert np.array_equal(res[0], res[1])


def test_uneven_launch_forms_agree():
    # one launch over all lengths, one launch per length and one launch per DM run the same unit
    # algorithm; the one-length kernel prefetches tiles and may be compiled with different FMA
    # contractions, so the series agree to FP32 rounding (rel-L2 < 1e-6) and the fold counts exactly
    design = ms.UnevenDesign(BW, F0, [0.0, 15.0, 120.0, 400.0, 900.0, 60.0])
    calls, nbins, period = 2, 64, 0.0011
    res = []
    for mode in (ms.LAUNCH_PERSISTENT, ms.LAUNCH_PER_CLASS, ms.LAUNCH_PER_DM):
        multi = ms.Multi(design, push_blocks=2)
        multi.set_launch_mode(mode)
        multi.fold(period, BW, nbins)
        n = multi.push_samples * calls
        raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 400.0, period, seed=34)
        outs = []
        for c in range(calls):
            multi.push(raw[c * multi.push_samples:(c + 1) * multi.push_samples])
            outs.append(np.stack([multi.output(i) for i in range(multi.D)]))
        _, sums, cnts = multi.fold_read()
        print(f"mode {mode}: classes {multi.num_classes}", flush=True)
        res.append((np.concatenate(outs, axis=1), sums, cnts))
        multi.close()
    for r in res[1:]:
        diff = np.abs(res[0][0] - r[0])
        print(f"  differing samples {np.mean(diff > 0):.2e}, max |diff| {diff.max():.2e}, "
              f"rel-L2 {rel_l2(r[0], res[0][0]):.2e}", flush=True)
        assert rel_l2(r[0], res[0][0]) < 1e-6
        assert np.array_equal(res[0][2], r[2])
        assert np.allclose(res[0][1], r[1], rtol=1e-6, atol=0)


def test_graph_push_replays_match_eager():
    # the push as one hipGraph (built on the first push, re-parameterised after): every replay
    # equals the eager one-launch push bit for bit, with host copies and syncs between replays
    design = ms.UnevenDesign(BW, F0, [0.0, 15.0, 120.0, 400.0, 900.0])
    calls, nbins, period = 4, 64, 0.0011
    res = []
    for graph in (False, True):
        multi = ms.Multi(design, push_blocks=2)
        multi.fold(period, BW, nbins)
        multi.set_graph(graph)
        n = multi.push_samples * calls
        raw, _ = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 400.0, period, seed=35)
        outs = []
        for c in range(calls):
            multi.push(raw[c * multi.push_samples:(c + 1) * multi.push_samples])
            outs.append(np.stack([multi.output(i) for i in range(multi.D)]))
            multi.synchronize()
        _, sums, cnts = multi.fold_read()
        print(f"graph={graph}: pushed {calls}", flush=True)
        res.append((outs, sums, cnts))
        multi.close()
    for c in range(calls):
        assert np.array_equal(res[0][0][c], res[1][0][c]), f"push {c}"
    assert np.array_equal(res[0][2], res[1][2])
    assert np.allclose(res[0][1], res[1][1], rtol=1e-12, atol=0)


def test_tuned_grid_at_the_bench_shape_matches_oracle():
    # the uneven bench's shape: band 784-816 MHz, pushes of 32 blocks of the longest length (2^22
    # samples), complex64 output, every length class's schedule tuned by msosm_multi_autotune (which
    # leaves the grid at the stream start); one DM of each class 2^14..2^18 against the FP64 oracle
    # over two pushes, in the one-launch form and in the one-launch-per-class form
    import torch

    bw, f0, period = 32e6, 784e6, 0.033
    design = ms.UnevenDesign(bw, f0, [4.0, 20.0, 80.0, 300.0, 900.0])
    assert sorted((2 * design.M).tolist()) == [1 << k for k in range(14, 19)]
    per, calls = 1 << 22, 2
    raw, x = make_input(ms.IN_C64, per * calls, bw, f0, 300.0, period, seed=35)
    dev = torch.from_numpy(raw.view(np.float32)).cuda()
    torch.cuda.synchronize()
    refs = []
    for i in range(len(design.dms)):
        M, dm = int(design.M[i]), float(design.dms[i])
        d, _ = orc.segmentation(bw, f0, dm, M)
        refs.append(orc.msosm(x, M, d[None], orc.transfer_function(bw, f0, dm, M)[None])[0])
    for mode in (ms.LAUNCH_PERSISTENT, ms.LAUNCH_PER_CLASS):
        multi = ms.Multi(design, push_blocks=32, out_flags=ms.OUT_C64, device=0)
        assert multi.push_samples == per
        best, tms = multi.autotune(dev.data_ptr(), calls)
        multi.set_launch_mode(mode)
        print(f"case tuned uneven grid at 32 blocks per push, launch mode {mode} (tuner's pick {best}, "
              f"{tms:.2f} ms per push)", flush=True)
        outs = [[] for _ in range(multi.D)]
        for c in range(calls):
            multi.push(dev.data_ptr() + c * per * 8)
            for i in range(multi.D):
                outs[i].append(multi.output(i))
        multi.close()
        for i, ref in enumerate(refs):
            err = rel_l2(np.concatenate(outs[i]), ref)
            print(f"  DM {design.dms[i]} (L = {2 * int(design.M[i])}): rel-L2 {err:.2e}", flush=True)
            assert err < TOL
