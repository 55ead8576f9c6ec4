"""The reference-compatible C++ classes (MSOSM_GPU_BATCH, OSM_GPU_BATCH, Fold_GPU), driven by
tests/cpp/check_reference_api.cpp with the reference harnesses' call sequences, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc
from tests.util import check_u16_codes, make_input, oracle_design, rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "check_reference_api")
BW, F0 = 16e6, 1392e6


def _run(args, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (make -C tests/cpp)")
    print("driver:", " ".join(map(str, args)), flush=True)
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return r.stdout


def test_save_mode_uint16(tmp_path):
    dms = [12.5, 30.0, 50.0]
    batch, calls = 4, 3
    design = ms.Design(BW, F0, dms)
    n = batch * design.M * calls
    raw, x = make_input(ms.IN_U16, 1 << (n - 1).bit_length(), BW, F0, 50.0, 0.0011, seed=21)
    raw, x = raw[:n], x[:n]
    (tmp_path / "dms.txt").write_text("\n".join(str(d) for d in dms))
    raw.astype(np.uint32).tofile(tmp_path / "in.u16")
    _run(["save", BW, F0, batch, "dms.txt", "in.u16", "out.u16"], tmp_path)
    got = np.fromfile(tmp_path / "out.u16", dtype=np.uint32)
    per = batch * design.M
    got = got.reshape(calls, len(dms), per).transpose(1, 0, 2).reshape(len(dms), n)
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)
    frac, worst = check_u16_codes(got, ref)
    print(f"  codes differing (all at rounding ties) {frac:.2e}, max diff {worst}", flush=True)


def test_fold_mode_profile(tmp_path):
    dm, batch, calls, period, nbins = 40.0, 4, 3, 0.0011, 128
    design = ms.Design(BW, F0, [dm])
    n = batch * design.M * calls
    raw, x = make_input(ms.IN_U16, 1 << (n - 1).bit_length(), BW, F0, dm, period, seed=22)
    raw, x = raw[:n], x[:n]
    raw.astype(np.uint32).tofile(tmp_path / "in.u16")
    _run(["fold", BW, F0, batch, dm, period, nbins, "in.u16", "fold.txt"], tmp_path)
    prof = np.loadtxt(tmp_path / "fold.txt")
    delays, H = oracle_design(BW, F0, design.dms, design.M)
    ref = orc.msosm(x, design.M, delays, H)[0]
    per = batch * design.M
    rs, rc = np.zeros(nbins), np.zeros(nbins, dtype=np.uint64)
    t0, dt = 0.0, 1.0 / float(np.float32(BW))
    for c in range(calls):
        orc.fold(np.abs(ref[c * per:(c + 1) * per]) ** 2, nbins, t0, dt, float(np.float32(period)), rs, rc)
        t0 += float(per) * dt
    np.testing.assert_allclose(prof, rs / rc, rtol=1e-5)


def test_plain_osm(tmp_path):
    dm, batch, calls = 50.0, 1, 4
    nd = orc.nd(BW, F0, dm)
    M = orc.osm_min_order(nd)  # 2^17: FFT length 2^18
    n = batch * M * calls
    raw, x = make_input(ms.IN_U16, n, BW, F0, dm, 0.033, seed=23)
    raw.astype(np.uint32).tofile(tmp_path / "in.u16")
    _run(["osm", BW, F0, batch, dm, "in.u16", "out.c64"], tmp_path)
    got = np.fromfile(tmp_path / "out.c64", dtype=np.complex64)
    H = orc.transfer_function(BW, F0, dm, M)[None, :]
    ref = orc.msosm(x, M, np.zeros((1, 2 * M), np.int32), H)[0]
    err = rel_l2(got, ref)
    print(f"OSM L=2^18 rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("writer", ["uring", "pwrite"])
def test_data_save_uring_files(tmp_path, writer):
    # MSOSM_DataSave_Uring with the reference harness sequence (msosm_save_simulated.cpp:124-149):
    # each DM's file holds its uint16 series, push after push, equal to the plan's own uint16 output
    dms = [12.5, 30.0, 50.0]
    batch, calls = 4, 5
    design = ms.Design(BW, F0, dms)
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_U16, 1 << (n - 1).bit_length(), BW, F0, 50.0, 0.0011, seed=24)
    raw = raw[:n]
    (tmp_path / "dms.txt").write_text("\n".join(str(d) for d in dms))
    raw.astype(np.uint32).tofile(tmp_path / "in.u16")
    out = tmp_path / "out"
    out.mkdir()
    _run(["uring", BW, F0, batch, "dms.txt", "in.u16", str(out), writer], tmp_path)
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_U16, out_flags=ms.OUT_U16)
    ref = []
    for c in range(calls):
        plan.push(raw[c * per:(c + 1) * per])
        ref.append(plan.output_u16())
    plan.close()
    ref = np.concatenate(ref, axis=1)
    for i, dm in enumerate(dms):
        path = ms.saver_file_name(str(out), "simulated", dms, i)
        got = np.fromfile(path, dtype=np.uint32)
        print(f"  {os.path.basename(path)}: {got.size} samples", flush=True)
        assert np.array_equal(got, ref[i])


def test_uneven_class_matches_per_dm_oracle(tmp_path):
    # MSOSM_GPU_UNEVEN (C++): every DM at its own length, driven like check_msosm.cpp
    dms = [5.0, 150.0, 600.0]
    design = ms.UnevenDesign(BW, F0, dms)
    push = 2 * int(design.M.max())
    calls = 3
    n = push * calls
    raw, x = make_input(ms.IN_C64, 1 << (n - 1).bit_length(), BW, F0, 150.0, 0.0011, seed=25)
    raw, x = raw[:n], x[:n]
    (tmp_path / "dms.txt").write_text("\n".join(str(d) for d in dms))
    raw.astype(np.complex64).tofile(tmp_path / "in.c64")
    _run(["uneven", BW, F0, 2, "dms.txt", "in.c64", "out.c64"], tmp_path)
    got = np.fromfile(tmp_path / "out.c64", dtype=np.complex64).reshape(len(dms), n)
    for i, dm in enumerate(dms):
        M = int(design.M[i])
        d, _ = orc.segmentation(BW, F0, float(np.float32(dm)), M)
        H = orc.transfer_function(BW, F0, float(np.float32(dm)), M)
        ref = orc.msosm(x, M, d[None], H[None])[0]
        err = rel_l2(got[i], ref)
        print(f"  DM {dm} (M={M}): rel-L2 {err:.2e}", flush=True)
        assert err < 1e-5


@pytest.mark.parametrize("npol", [2, 1])
def test_psrdada_dual_pol_fold(tmp_path, npol):
    # run_psrdada_GPU_batch (msosm_psrdada.cpp:88-153) on a synthetic DADA file: a leading read of
    # code-0 samples is skipped, each polarisation is dedispersed, and sqrt(|a|^2 + |b|^2) (one
    # polarisation: |a|^2) is folded (gpu_kernel.cu:430-444, fold_gpu.cpp:48-54)
    from tests.util import write_dada

    dm, period, count, nbins = 30.0, 0.0011, 4, 64
    bw_mhz, fc_mhz = 16.0, 1400.0
    f0 = (fc_mhz - bw_mhz / 2) * 1e6
    M = ms.Design(bw_mhz * 1e6, f0, [dm]).M
    per = count * M
    calls = 3
    n2 = 1 << (per * calls - 1).bit_length()
    raw1, x1 = make_input(ms.IN_U16, n2, bw_mhz * 1e6, f0, dm, period, seed=26)
    raw2, x2 = make_input(ms.IN_U16, n2, bw_mhz * 1e6, f0, dm, period, seed=27)
    raw1, x1, raw2, x2 = raw1[:per * calls], x1[:per * calls], raw2[:per * calls], x2[:per * calls]
    lead = np.zeros(per, dtype=np.uint32)
    path = tmp_path / "psr.dada"
    write_dada(path, np.concatenate([lead, raw1]), np.concatenate([lead, raw2]) if npol == 2 else None,
               bw_mhz, fc_mhz)
    out = _run(["psrdada", str(path), dm, period, count, nbins], tmp_path)
    assert f"calls {calls} of {calls + 1}" in out
    prof = np.loadtxt(tmp_path / "psr_dedisp_folded.txt")
    delays, H = oracle_design(bw_mhz * 1e6, f0, [dm], M)
    y1 = orc.msosm(x1, M, delays, H)[0]
    I = np.abs(y1) ** 2
    if npol == 2:
        y2 = orc.msosm(x2, M, delays, H)[0]
        I = np.sqrt(np.abs(y1) ** 2 + np.abs(y2) ** 2)
    sums, cnts = np.zeros(nbins), np.zeros(nbins, dtype=np.uint64)
    dt = 1.0 / float(np.float32(bw_mhz * 1e6))
    t0 = 0.0
    for c in range(calls):
        orc.fold(I[c * per:(c + 1) * per], nbins, t0, dt, float(np.float32(period)), sums, cnts)
        t0 += float(per) * dt
    ref = np.where(cnts > 0, sums / np.maximum(cnts, 1), 0.0)
    err = np.max(np.abs(prof - ref) / np.abs(ref))
    print(f"  npol {npol}: profile max rel err {err:.2e}", flush=True)
    assert err < 1e-5


def test_simulated_complex_signal_class(tmp_path):
    # SimulatedComplexSignal(bw, dm, f0, period, "uint16").generate_pulsar_signal_block(dummy, valid), as
    # check_msosm.cpp:128-133 calls it: the same record, bit for bit, as the C ABI's simulator
    bw, dm, f0, period, dummy, valid = 16e6, 30.0, 1392e6, 1e-3, 2, 130
    out = tmp_path / "sim.u16"
    text = _run(["simulate", bw, dm, f0, period, dummy, valid, out], tmp_path)
    p = ms.simulate_params(bw, dm, f0, period, dummy, valid)
    assert f"Np {p.Np} Nd {p.Nd} range {p.range} size {p.samples}" in text
    got = np.fromfile(out, dtype=np.uint32)
    assert np.array_equal(got, ms.simulate_pulsar_block(bw, dm, f0, period, dummy, valid))


def test_simulated_complex_signal_new(tmp_path):
    # generate_pulsar_signal_new(repeat) in "uint16" mode fills both signal and signal_u16
    bw, dm, f0, period, repeat = 16e6, 30.0, 1392e6, 1e-3, 12
    out, outc = tmp_path / "sim.u16", tmp_path / "sim.c64"
    text = _run(["simulate_new", bw, dm, f0, period, repeat, out, outc], tmp_path)
    p = ms.simulate_params_new(bw, dm, f0, period, repeat)
    assert f"Np {p.Np} Nd {p.Nd} range {p.range} size {p.samples}" in text
    codes, series = ms.simulate_pulsar_new(bw, dm, f0, period, repeat, series=True)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), codes)
    assert np.array_equal(np.fromfile(outc, dtype=np.complex64), series)


@pytest.mark.parametrize("pinned", [0, 1])
def test_simulated_complex_signal_pulse(tmp_path, pinned):
    # generate_pulsar_signal(repeat, false, 0, pinned), as msosm_save_simulated.cpp:115 calls it
    bw, dm, f0, period, repeat = 16e6, 30.0, 1392e6, 1e-3, 5
    out, outc = tmp_path / "sim.u16", tmp_path / "sim.c64"
    text = _run(["simulate_pulse", bw, dm, f0, period, repeat, pinned, out, outc], tmp_path)
    p = ms.simulate_params_pulse(bw, dm, f0, period, repeat)
    assert f"Np {p.Np} Nd {p.Nd} range {p.range} size {p.samples}" in text
    codes, series = ms.simulate_pulsar(bw, dm, f0, period, repeat, series=True)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), codes)
    assert np.array_equal(np.fromfile(outc, dtype=np.complex64), series)

