"""Host side of the GPU pulsar simulator (simulated_complex_signal.cpp:16-41, 43-72, 74-198), on the CPU.

The noise comes from the standard library's mt19937 and normal_distribution, as in the reference, so
it is pinned here by the invariants of the reference's construction: the seed-2 periods do not depend
on how many seed-1 periods precede them, the seed-1 prefix is shared, the noise power follows the
three-Gaussian profile (peak power 10^8) and the off-pulse samples underflow to exactly zero. The
chirp is pinned by |H| = 1, H = 1 at zero frequency offset, and its phase slope, which must equal
the physical group delay kdm DM (f0^-2 - (f0 + f)^-2).
"""
import numpy as np
import pytest

import msosm_amd as ms

BAND = (16e6, 30.0, 1392e6, 1e-3)  # bw (= fs), DM, f0, period


def test_params_follow_the_reference_definitions():
    for bw, dm, f0, period in [BAND, (16e6, 30.0, 1392e6, 4e-3), (64e6, 100.0, 1218e6, 5e-3), (32e6, 0.0, 784e6, 2e-3)]:
        for dummy, valid in [(0, 128), (3, 200)]:
            p = ms.simulate_params(bw, dm, f0, period, dummy, valid)
            assert p.Np == int(np.float32(period) * np.float32(bw))
            assert p.Nd == max(0, ms.design_nd(bw, f0, dm))  # the constructor's Nd is Prepare_MSOSM's
            assert p.range == -(-p.Nd // p.Np)
            long_size = (2 * (p.range + 2) + 128) * p.Np
            assert p.fftpoint >= long_size and p.fftpoint < 2 * long_size and p.fftpoint & (p.fftpoint - 1) == 0
            assert p.start == p.fftpoint - long_size + (p.range + 2) * p.Np
            assert p.samples == (dummy + valid) * p.Np
            assert p.chunks == -(-(dummy + valid) // 128)
            assert p.noise_len == p.fftpoint + (p.chunks - 1) * 128 * p.Np


def test_noise_seed_structure():
    bw, dm, f0, period = BAND
    a = ms.simulate_noise(bw, dm, f0, period, 0, 128)
    b = ms.simulate_noise(bw, dm, f0, period, 2, 128)
    pa, pb = ms.simulate_params(bw, dm, f0, period, 0, 128), ms.simulate_params(bw, dm, f0, period, 2, 128)
    Np, guard = pa.Np, pa.range + 2
    off_a, off_b = pa.fftpoint - (2 * guard + 128) * Np, pb.fftpoint - (2 * guard + 128) * Np
    # seed 1: the first guard - 1 periods are the same stream in both
    assert np.array_equal(a[off_a:off_a + (guard - 1) * Np], b[off_b:off_b + (guard - 1) * Np])
    # seed 2 restarts at the valid periods: identical whatever came before
    va = a[off_a + (guard - 1) * Np:off_a + (guard - 1 + 128) * Np]
    vb = b[off_b + (guard + 1) * Np:off_b + (guard + 1 + 128) * Np]
    assert np.array_equal(va, vb)
    # and differs from the seed-1 stream
    assert not np.array_equal(va[:Np], a[off_a:off_a + Np])
    # nothing before the first period
    assert not np.any(a[:off_a])


def test_noise_follows_the_three_gaussian_profile():
    bw, dm, f0, period = BAND
    p = ms.simulate_params(bw, dm, f0, period, 0, 128)
    a = ms.simulate_noise(bw, dm, f0, period, 0, 128)
    off = p.fftpoint - (2 * (p.range + 2) + 128) * p.Np
    per = a[off:off + (2 * (p.range + 2) + 128) * p.Np].reshape(-1, p.Np)
    x = np.arange(p.Np) / p.Np
    prof = (0.45 * np.exp(-0.5 * ((x - 0.455) / 0.008) ** 2) + np.exp(-0.5 * ((x - 0.5) / 0.012) ** 2)
            + 0.28 * np.exp(-0.5 * ((x - 0.555) / 0.006) ** 2))
    expect = (prof * 1e4 / prof.max()) ** 2  # E|x|^2: two components of variance (profile scale)^2 / 2
    power = np.mean(np.abs(per.astype(np.complex128)) ** 2, axis=0)
    on = expect > 1e6
    ratio = power[on].sum() / expect[on].sum()
    print(f"noise power / profile power on-pulse: {ratio:.4f}")
    assert abs(ratio - 1.0) < 0.02
    # far from the pulse the amplitudes underflow in float: exactly zero
    assert not np.any(per[:, x < 0.3])


def test_chirp_modulus_and_group_delay():
    bw, dm, f0, _ = BAND
    n = 1 << 16
    h = ms.simulate_chirp(bw, dm, f0, n).astype(np.complex128)
    assert np.max(np.abs(np.abs(h) - 1.0)) < 1e-6
    assert h[n // 2] == 1.0  # f_disp = 0
    # bins k < n/2 are the upper half of the band (f = fs/2 + k fs/n), k >= n/2 the lower half
    k = np.arange(n)
    f = np.where(k < n // 2, bw / 2 + k * bw / n, k * bw / n - bw / 2)
    order = np.argsort(f)
    phase = np.unwrap(np.angle(h[order]))
    tau = 4.15e15 * dm * (1 / f0 ** 2 - 1 / (f0 + f[order]) ** 2)
    slope = np.diff(phase) / (2 * np.pi * bw / n)  # d phase / d (2 pi f)
    mid = 0.5 * (tau[1:] + tau[:-1])
    rel = np.max(np.abs(slope - mid)) / tau.max()
    print(f"group delay from the chirp phase vs kdm DM (f0^-2 - (f0+f)^-2): max rel {rel:.2e}")
    assert rel < 1e-3


def test_unsupported_and_bad_arguments_raise():
    with pytest.raises(ms.MsosmError):
        ms.simulate_params(0.0, 1.0, 1e9, 1e-3)
    with pytest.raises(ms.MsosmError):
        ms.simulate_params(16e6, 1.0, 1e9, 1e-3, 0, 0)


# generate_pulsar_signal_new (simulated_complex_signal.cpp:279-467)

def test_params_new_follow_the_reference_definitions():
    for bw, dm, f0, period in [BAND, (64e6, 100.0, 1218e6, 5e-3), (32e6, 0.0, 784e6, 2e-3), (128e6, 75.0, 1e9, 0.002048)]:
        for repeat in (1, 16, 200):
            p = ms.simulate_params_new(bw, dm, f0, period, repeat)
            b = ms.simulate_params(bw, dm, f0, period, 0, repeat)
            assert (p.Np, p.Nd, p.range) == (b.Np, b.Nd, b.range)  # the constructor's, shared
            guard = p.range + 2
            assert p.periods == repeat + 2 * guard
            need = p.periods * p.Np + p.Nd
            assert p.fftpoint >= need and p.fftpoint < 2 * need and p.fftpoint & (p.fftpoint - 1) == 0
            assert p.filter_point >= p.Nd + 1 and p.filter_point & (p.filter_point - 1) == 0
            assert p.filter_point == 1 or p.filter_point < 2 * (p.Nd + 1)
            assert p.start == guard * p.Np + p.Nd and p.samples == repeat * p.Np
            assert p.start + p.samples <= p.fftpoint


def test_noise_new_is_one_seed1_stream():
    bw, dm, f0, period = BAND
    p = ms.simulate_params_new(bw, dm, f0, period, 4)
    rec = ms.simulate_noise_new(bw, dm, f0, period, 4)
    guard = p.range + 2
    # the block generator draws its first guard + dummy - 1 periods from the same seed-1 stream
    dummy = p.periods - guard + 1
    b = ms.simulate_params(bw, dm, f0, period, dummy, 1)
    blk = ms.simulate_noise(bw, dm, f0, period, dummy, 1)
    off = b.fftpoint - (2 * guard + 128) * b.Np
    assert np.array_equal(rec[:p.periods * p.Np], blk[off:off + p.periods * p.Np])
    assert not np.any(rec[p.periods * p.Np:])  # zero padding up to fftpoint
    assert np.any(rec[(p.periods - 1) * p.Np:p.periods * p.Np])


def test_new_bad_arguments_raise():
    with pytest.raises(ms.MsosmError):
        ms.simulate_params_new(16e6, 1.0, 1e9, 1e-3, 0)
    with pytest.raises(ms.MsosmError):
        ms.simulate_params_new(16e6, -1.0, 1e9, 1e-3, 4)


# generate_pulsar_signal (simulated_complex_signal.cpp:469-706)

def test_params_pulse_follow_the_reference_definitions():
    for bw, dm, f0, period in [BAND, (64e6, 100.0, 1218e6, 5e-3), (128e6, 75.0, 1e9, 0.002048), (16e6, 1.0, 1392e6, 1e-4)]:
        p = ms.simulate_params_pulse(bw, dm, f0, period, 3)
        b = ms.simulate_params(bw, dm, f0, period, 0, 3)
        assert (p.Np, p.Nd) == (b.Np, b.Nd)
        assert p.range == int(np.ceil(np.float32(p.Nd) / np.float32(p.Np)))  # float division there
        assert p.order == 1 << int(np.ceil(np.log2(float(p.range * p.Np))))
        assert p.fftpoint == 2 * p.order and p.samples == 3 * p.Np


def test_pulse_rejects_zero_dm_and_noise():
    with pytest.raises(ms.MsosmError):
        ms.simulate_params_pulse(16e6, 0.0, 1392e6, 1e-3, 1)  # no dispersion: order undefined there
    # the reference's add_noise branch races on period 0: refused before any device work
    with pytest.raises(ms.MsosmError) as e:
        ms.simulate_pulsar(16e6, 30.0, 1392e6, 1e-3, 2, add_noise=True, snr=-10.0)
    assert e.value.code == ms.ERR_UNSUPPORTED


def test_host_alloc_round_trip():
    # page-locked host memory needs the HIP runtime but no device work; bytes = 0 gives NULL
    p = ms.C.c_void_p()
    assert ms.lib().msosm_host_alloc(0, ms.C.byref(p)) == ms.OK and not p.value
    assert ms.lib().msosm_host_free(None) == ms.OK


@pytest.mark.parametrize("gen", ["block", "new", "pulse"])
def test_unsupported_fft_length_is_refused_before_device_work(gen):
    # 2^23 needs a 2^11 factor, which the in-workgroup engine lacks; the check precedes any HIP call
    if gen == "block":
        p = ms.simulate_params(16e6, 30.0, 1392e6, 2e-3, 0, 1)
        call = lambda: ms.simulate_pulsar_block(16e6, 30.0, 1392e6, 2e-3, 0, 1)
    elif gen == "new":
        p = ms.simulate_params_new(16e6, 30.0, 1392e6, 1e-3, 300)
        call = lambda: ms.simulate_pulsar_new(16e6, 30.0, 1392e6, 1e-3, 300)
    else:
        p = ms.simulate_params_pulse(16e6, 30.0, 1392e6, 0.2, 1)
        call = lambda: ms.simulate_pulsar(16e6, 30.0, 1392e6, 0.2, 1)
    assert p.fftpoint == 1 << 23
    with pytest.raises(ms.MsosmError) as e:
        call()
    assert e.value.code == ms.ERR_UNSUPPORTED
