"""Pins of the CPU oracle (no GPU): every oracle function is checked against numpy, closed forms,
invariants or brute force, so that a dropped term, a sign error or a transposed index fails here.

The reference ships no golden vectors, fixtures or tests for this path (its checks are visual), so
these are the pins; tests/golden/ holds small fixtures derived from the closed forms below.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

# (bw, fc, dm) of the five benchmark configurations and the reference's check default
CONFIGS = {
    "cfg1": (16e6, 1400e6, 50.0),
    "cfg2": (64e6, 1250e6, 57.3),
    "cfg2_max": (64e6, 1250e6, 100.0),
    "cfg3_max": (32e6, 800e6, 1000.0),
    "cfg4_max": (50e6, 350e6, 1500.0),
    "cfg5": (16e6, 1400e6, 123.456),
    "ref_default": (128e6, 1064e6, 75.0),
}
# Nd, MS-OSM M, delaycount and OSM M for each configuration, from the reference formulas
# (msosm_prepare.cpp:33,60-61,217; osm_prepare.cpp:21-23), an independent evaluation
EXPECTED = {
    "cfg1": (38719, 8192, 5, 131072),
    "cfg2": (998693, 32768, 31, 2097152),
    "cfg2_max": (1742919, 32768, 54, 4194304),
    "cfg3_max": (16613287, 131072, 127, 33554432),
    "cfg4_max": (733412230, 1048576, 700, 2147483648),
    "cfg5": (95603, 8192, 12, 262144),
    "ref_default": (8528697, 131072, 66, 33554432),
}


# --------------------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 4, 8, 64, 1024, 1 << 14])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(orc.fft(x, -1) - ref)) / scale < 1e-13
    assert np.max(np.abs(orc.fft(x, +1) - np.fft.ifft(x) * n)) / scale < 1e-13


def test_fft_closed_forms():
    n = 4096
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.allclose(orc.fft(d), 1.0, atol=0, rtol=1e-15)
    k0 = 37
    e = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    X = orc.fft(e, -1)  # e^{+2 pi i k0 t/n} -> n * delta[k - k0]
    expect = np.zeros(n, complex)
    expect[k0] = n
    assert np.max(np.abs(X - expect)) < 1e-9
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = orc.fft(x)
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) / np.sum(np.abs(x) ** 2) < 1e-13
    assert np.max(np.abs(orc.fft(X, +1) / n - x)) < 1e-13
    with pytest.raises(ValueError):
        orc.fft(np.zeros(12, complex))


# ------------------------------------------------------------------------------------ design
@pytest.mark.parametrize("name", list(CONFIGS))
def test_design_numbers(name):
    bw, fc, dm = CONFIGS[name]
    f0 = fc - bw / 2
    nd = orc.nd(bw, f0, dm)
    M = orc.min_order(nd)
    assert (nd, M, orc.delaycount(nd, M), orc.osm_min_order(nd)) == EXPECTED[name]


@pytest.mark.parametrize("name", list(CONFIGS))
def test_nd_closed_form(name):
    # Nd = floor(kdm DM (f0^-2 - (f0+bw)^-2) fs) in exact rational arithmetic, with the float
    # rounding of kdm*DM and 1/fs the reference applies; the oracle must agree to the sample
    bw, fc, dm = CONFIGS[name]
    f0 = np.float32(fc - bw / 2)
    bwf = np.float32(bw)
    kdm_dm = np.float32(np.float32(4.15e15) * np.float32(dm))
    inv_fs = np.float32(np.float32(1) / bwf)
    f0b = np.float32(f0 + bwf)
    exact = Fraction(float(kdm_dm)) * (Fraction(1) / Fraction(float(f0)) ** 2 - Fraction(1) / Fraction(float(f0b)) ** 2)
    exact /= Fraction(float(inv_fs))
    nd = orc.nd(bw, float(f0), dm)
    assert abs(nd - math.floor(exact)) <= 1


def _tau(f_off, bw, f0, dm):
    """group delay [s] of a tone f_off above the band bottom, kdm DM (f0^-2 - (f0+f)^-2), with the
    dispersion constant and DM as the float values the reference stores (msosm_prepare.h:37-39)"""
    kdm = float(np.float32(4.15e15))
    return kdm * float(np.float32(dm)) * (1.0 / f0 ** 2 - 1.0 / (f0 + f_off) ** 2)


def _bin_offset(k, M, fs):
    """frequency above the band bottom of FFT bin k of a 2M-point transform at rate fs"""
    L = 2 * M
    f = k * fs / L if k < M else (k - L) * fs / L  # signed baseband frequency
    return f + fs / 2


@pytest.mark.parametrize("name", ["cfg1", "cfg2", "cfg2_max", "cfg3_max", "cfg5", "ref_default"])
def test_segmentation_invariants(name):
    bw, fc, dm = CONFIGS[name]
    f0 = fc - bw / 2
    nd = orc.nd(bw, f0, dm)
    M = orc.min_order(nd)
    d, dc = orc.segmentation(bw, f0, dm, M)
    L = 2 * M
    assert dc == orc.delaycount(nd, M)
    assert d.min() == 0 and d.max() == dc - 1
    assert d[M] == 0  # band bottom is not delayed
    assert d[M - 1] == dc - 1  # band top gets the largest delay
    natural = np.concatenate([d[M:], d[:M]])  # bottom -> top
    assert np.all(np.diff(natural) >= 0)
    assert np.all(np.diff(natural) <= 1)
    # each boundary is the floor of the FP64 solution of tau(f) = j M / fs on the bin grid: the
    # first bin with delay j lies 0..1 bins below the exact boundary (+- float rounding of w_j)
    block = M / bw
    kdm = float(np.float32(4.15e15)) * float(np.float32(dm))
    width = bw / L
    for j in range(1, dc):
        idx = np.nonzero(natural >= j)[0][0]  # first bin (from the bottom) with delay >= j
        f_first = _bin_offset((idx + M) % L, M, bw)
        f_exact = 1.0 / math.sqrt(1.0 / f0 ** 2 - j * block / kdm) - f0
        assert -0.05 <= (f_exact - f_first) / width <= 1.05
        assert _tau(f_first - width, bw, f0, dm) < j * block


def test_segmentation_special_cases():
    # DM 0: defined as no delay at all
    d, dc = orc.segmentation(16e6, 1392e6, 0.0, 4096)
    assert dc == 1 and not d.any()
    # dc == 1 (smear shorter than a block): every bin undelayed
    nd = orc.nd(16e6, 1392e6, 0.5)
    assert orc.delaycount(nd, 8192) == 1
    d, dc = orc.segmentation(16e6, 1392e6, 0.5, 8192)
    assert dc == 1 and not d.any()


def test_build_segments():
    d, dc = orc.segmentation(64e6, 1218e6, 100.0, 32768)
    segs = orc.build_segments(d)
    starts, lens, groups = segs[:, 0], segs[:, 1], segs[:, 2]
    assert starts[0] == 0 and np.all(starts[1:] == starts[:-1] + lens[:-1]) and starts[-1] + lens[-1] == d.size
    assert lens.max() <= 256 and lens.min() >= 1
    for s, l, g in segs:
        assert np.all(d[s:s + l] == g)
    # maximal runs: a new segment starts only on a delay change or after 256 bins
    for i in range(1, len(segs)):
        assert groups[i] != groups[i - 1] or lens[i - 1] == 256


@pytest.mark.parametrize("name", ["cfg1", "cfg2_max", "cfg5"])
def test_transfer_function_closed_forms(name):
    bw, fc, dm = CONFIGS[name]
    f0 = fc - bw / 2
    nd = orc.nd(bw, f0, dm)
    M = orc.min_order(nd)
    H = orc.transfer_function(bw, f0, dm, M)
    L = 2 * M
    # even bins are the chirp itself: c_n = e^{-i phi(f_n)}, f_n = ((n + M/2) mod M) fs / M
    n = np.arange(M)
    fn = ((n + M // 2) % M) * float(np.float32(bw)) / M
    f0d = float(np.float32(f0))
    phi = 2 * np.pi * 4.15e15 * float(np.float32(dm)) * fn ** 2 / (f0d ** 2 * (f0d + fn))
    assert np.max(np.abs(H[0::2] - np.exp(-1j * phi))) < 1e-6 * max(1.0, phi.max() / 1e4)
    assert np.max(np.abs(np.abs(H[0::2]) - 1)) < 1e-12
    # the impulse response has M taps
    h = np.fft.ifft(H)
    assert np.max(np.abs(h[M:])) < 1e-12
    # DM 0 gives the identity
    H0 = orc.transfer_function(bw, f0, 0.0, M)
    assert np.max(np.abs(H0 - 1)) < 1e-12


def test_transfer_function_fp32_phase_is_worse():
    bw, f0, dm, M = 64e6, 1218e6, 100.0, 32768
    H64 = orc.transfer_function(bw, f0, dm, M)
    H32 = orc.transfer_function(bw, f0, dm, M, phase_fp32=True)
    err = np.max(np.abs(H32[0::2] - H64[0::2]))
    assert err > 1e-3  # the float-w phase loses whole radians at 1e5-1e6 cycles


# ------------------------------------------------------------------------------ MS-OSM / OSM
def test_osm_equals_causal_convolution():
    # delays all 0: overlap-save with an M-tap FIR is exactly the causal linear convolution
    rng = np.random.default_rng(2)
    M, N = 64, 4096
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    h = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    H = np.fft.fft(np.concatenate([h, np.zeros(M)]))
    out = orc.msosm(x, M, np.zeros((1, 2 * M), np.int32), H[None, :])[0]
    direct = np.convolve(x, h)[:N]
    assert np.max(np.abs(out - direct)) / np.max(np.abs(direct)) < 1e-13


def test_msosm_identity_and_pure_delays():
    rng = np.random.default_rng(3)
    M, N = 128, 128 * 20
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    L = 2 * M
    ones = np.ones((1, L), complex)
    # identity
    out = orc.msosm(x, M, np.zeros((1, L), np.int32), ones)[0]
    assert np.max(np.abs(out - x)) < 1e-13
    # a whole-block delay on every bin delays the output by M samples
    out = orc.msosm(x, M, np.ones((1, L), np.int32), ones)[0]
    assert np.max(np.abs(out[M:] - x[:-M])) < 1e-13 and np.max(np.abs(out[:M])) < 1e-13
    # a sub-block delay of d samples through H = e^{-2 pi i k d / L}
    dsm = 17
    Hd = np.exp(-2j * np.pi * np.arange(L) * dsm / L)[None, :]
    out = orc.msosm(x, M, np.zeros((1, L), np.int32), Hd)[0]
    assert np.max(np.abs(out[dsm:] - x[:-dsm])) < 1e-12
    # delaying a band by one block: only that band's content moves (linearity across segments)
    dl = np.zeros((1, L), np.int32)
    dl[0, : M // 2] = 1
    out = orc.msosm(x, M, dl, ones)[0]
    part = orc.msosm(x, M, np.ones((1, L), np.int32), ones)[0] - orc.msosm(x, M, np.zeros((1, L), np.int32), ones)[0]
    assert np.linalg.norm(out - x) > 0.1 * np.linalg.norm(part)


def test_msosm_block_shift_covariance():
    bw, f0, dm = 16e6, 1392e6, 50.0
    M = orc.min_order(orc.nd(bw, f0, dm))
    d, _ = orc.segmentation(bw, f0, dm, M)
    H = orc.transfer_function(bw, f0, dm, M)
    rng = np.random.default_rng(4)
    N = 16 * M
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    a = orc.msosm(x, M, d[None], H[None])[0]
    xs = np.concatenate([np.zeros(M), x[:-M]])
    b = orc.msosm(xs, M, d[None], H[None])[0]
    assert np.max(np.abs(b[M:] - a[:-M])) < 1e-12 * np.max(np.abs(a))


def test_msosm_dedisperses_the_simulated_pulse():
    # physics: a narrow pulse dispersed with the direct chirp over the whole record is restored
    bw, f0, dm = 16e6, 1392e6, 50.0
    M = orc.min_order(orc.nd(bw, f0, dm))
    d, _ = orc.segmentation(bw, f0, dm, M)
    H = orc.transfer_function(bw, f0, dm, M)
    n = 1 << 18
    period = n / bw / 2  # two pulses in the record
    x = orc.pulsar(n, bw, f0, dm, period, duty=0.0005, amp=20.0, sigma=0.0, seed=1)
    out = orc.msosm(x, M, d[None], H[None])[0]
    p = np.abs(out) ** 2
    raw = np.abs(orc.pulsar(n, bw, f0, 0.0, period, duty=0.0005, amp=20.0, sigma=0.0, seed=1)) ** 2
    # the pulse at phase 0.5 of period 2 is far from the record ends
    lo, hi = int(1.2 * period * bw), int(1.8 * period * bw)
    peak_out = lo + int(np.argmax(p[lo:hi]))
    peak_true = lo + int(np.argmax(raw[lo:hi]))
    assert abs(peak_out - peak_true) < 0.0005 * period * bw * 3
    # dispersion smears the pulse: the dedispersed peak stands far above the dispersed data's
    assert p[lo:hi].max() > 5 * (np.abs(x[lo:hi]) ** 2).max()


# -------------------------------------------------------------------------------------- fold
def _py_bin(i, nbins, t_start, dt, period):
    t = t_start + float(i) * dt  # Python floats are IEEE doubles and never contract
    phase = math.fmod(t / period, 1.0)
    b = int(math.floor(phase * float(nbins)))
    return min(b, nbins - 1)


def test_fold_bin_matches_ieee_evaluation():
    rng = np.random.default_rng(5)
    for _ in range(20):
        fs = float(np.float32(rng.uniform(1e6, 2e8)))
        period = float(np.float32(rng.uniform(1e-3, 5e-2)))
        nbins = int(rng.choice([64, 256, 1000, 1024, 2048]))
        t_start = float(rng.integers(0, 1 << 30)) * (1.0 / fs)
        dt = 1.0 / fs
        for i in rng.integers(0, 1 << 26, size=200):
            assert orc.fold_bin(int(i), nbins, t_start, dt, period) == _py_bin(int(i), nbins, t_start, dt, period)


def test_fold_bin_closed_form():
    # fs = 2^20 Hz, P = 2^-4 s: t/P = i / 2^16 exactly, so bin = (i mod 2^16) * nbins / 2^16
    fs, period, nbins = float(1 << 20), 1.0 / 16, 1024
    i = np.arange(0, 300000, 7)
    got = np.array([orc.fold_bin(int(k), nbins, 0.0, 1.0 / fs, period) for k in i])
    assert np.array_equal(got, ((i % 65536) * nbins) // 65536)


def test_fold_counts_and_zero_skip():
    size, nbins = 100000, 128
    I = np.ones(size)
    I[::10] = 0.0
    s, c, bins = orc.fold(I, nbins, 0.0, 1e-6, 0.00123, want_bins=True)
    assert int(c.sum()) == size - size // 10
    assert np.all(bins[::10] == -1)
    nz = c > 0
    assert np.allclose(s[nz] / c[nz], 1.0)


# ------------------------------------------------------------------------------------- codecs
def test_unpack_u16_exhaustive():
    codes = np.arange(65536, dtype=np.uint32)
    packed = (codes[::-1] << 16) | codes
    out = orc.unpack_u16(packed)
    expect_re = np.where(codes == 0, 0.0, codes.astype(float) - 32768.0)
    expect_im = np.where(codes[::-1] == 0, 0.0, codes[::-1].astype(float) - 32768.0)
    assert np.array_equal(out.real, expect_re) and np.array_equal(out.imag, expect_im)


@pytest.mark.parametrize("x,code", [(0.0, 0), (-0.0, 0), (1e-30, 32768), (0.5, 32768), (1.5, 32770),
                                    (2.5, 32770), (-32768.0, 0), (-40000.0, 0), (32766.6, 65535),
                                    (32767.0, 65535), (1e9, 65535), (float("nan"), 0), (-1.0, 32767)])
def test_encode_u16_edges(x, code):
    assert orc.encode_u16_pair(x, 0.0) & 0xFFFF == code
    assert orc.encode_u16_pair(0.0, x) >> 16 == code
