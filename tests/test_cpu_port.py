"""The FP32 CPU port (the timed CPU baseline of bench.py) against the FP64 oracle.

oracle/cpu_msosm.c restates the same MS-OSM + fold in FP32 with FFTW3-API transforms (Intel MKL's
FFTW3 interface when it loads, else its own radix-2 FFT). Both backends must reproduce the oracle's
series within FP32 rounding and its fold counts exactly (same bins, same zero skips).
"""
import numpy as np
import pytest

from oracle import oracle as orc

BW, F0, M, NBINS, PERIOD = 16e6, 1392e6, 1024, 256, 0.0011
DMS = [0.0, 5.0, 12.0]


@pytest.fixture(scope="module")
def case():
    n = 16 * M
    x = orc.pulsar(1 << (n - 1).bit_length(), BW, F0, 10.0, PERIOD, seed=3)[:n]
    delays = np.stack([orc.segmentation(BW, F0, float(np.float32(d)), M)[0] for d in DMS])
    H = np.stack([orc.transfer_function(BW, F0, float(np.float32(d)), M) for d in DMS])
    return x, delays, H, orc.msosm(x, M, delays, H)


@pytest.mark.parametrize("force_own", [False, True], ids=["fftw3-api", "own-fft"])
def test_cpu_port_matches_oracle(case, force_own):
    x, delays, H, ref = case
    if not force_own and not orc.cpu_backend(False).startswith("FFTW3 API"):
        pytest.skip("MKL's FFTW3 interface is not loadable here")
    sums, cnts, out = orc.cpu_msosm_fold(x, M, delays, H, NBINS, 0.0, 1.0 / BW, PERIOD, nthreads=3,
                                         force_own=force_own, want_out=True)
    for d in range(len(DMS)):
        err = np.linalg.norm(out[d] - ref[d]) / np.linalg.norm(ref[d])
        assert err < 2e-6, (DMS[d], err)  # FP32 transforms of 2^11 points: ~2e-7 measured
        rs, rc = np.zeros(NBINS), np.zeros(NBINS, dtype=np.uint64)
        orc.fold(np.abs(ref[d]) ** 2, NBINS, 0.0, 1.0 / BW, PERIOD, rs, rc)
        assert np.array_equal(cnts[d], rc)
        np.testing.assert_allclose(sums[d], rs, rtol=2e-5)


def test_cpu_port_is_thread_count_invariant(case):
    x, delays, H, _ = case
    a = orc.cpu_msosm_fold(x, M, delays, H, NBINS, 0.0, 1.0 / BW, PERIOD, nthreads=1, want_out=True)
    b = orc.cpu_msosm_fold(x, M, delays, H, NBINS, 0.0, 1.0 / BW, PERIOD, nthreads=4, want_out=True)
    # each DM is owned by one thread: identical arithmetic whatever the thread count
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_cpu_port_rejects_bad_shapes(case):
    x, delays, H, _ = case
    with pytest.raises(ValueError):
        orc.cpu_msosm_fold(x, M, delays[:, :M], H, NBINS, 0.0, 1.0 / BW, PERIOD)
    with pytest.raises(RuntimeError):
        orc.cpu_msosm_fold(x[: M + 1], M, delays, H, NBINS, 0.0, 1.0 / BW, PERIOD)


def test_cpu_port_folds_in_calls_like_fold_gpu(case):
    # calls of 4M samples: t_start advances by 4M*dt per call (fold_gpu.cpp:53), as the oracle's fold
    # is driven call by call in the GPU tests
    x, delays, H, ref = case
    per = 4 * M
    sums, cnts, _ = orc.cpu_msosm_fold(x, M, delays, H, NBINS, 0.0, 1.0 / BW, PERIOD, call_size=per)
    for d in range(len(DMS)):
        rs, rc = np.zeros(NBINS), np.zeros(NBINS, dtype=np.uint64)
        t0 = 0.0
        for c in range(x.size // per):
            orc.fold(np.abs(ref[d, c * per:(c + 1) * per]) ** 2, NBINS, t0, 1.0 / BW, PERIOD, rs, rc)
            t0 += float(per) * (1.0 / BW)
        assert np.array_equal(cnts[d], rc)
        np.testing.assert_allclose(sums[d], rs, rtol=2e-5)
