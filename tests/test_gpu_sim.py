"""GPU pulsar simulator (simulated_complex_signal.cpp:74-277, 279-467) against FP64 restatements.

The device disperses each chunk of 128 periods with one complex FFT of `fftpoint` points: one
workgroup per transform up to 2^22 points, a six-step decomposition over two in-workgroup lengths
above. The FP64 reference takes the same noise and chirp (pinned in test_sim.py) and does the
reference's steps with numpy: FFT, times the chirp, inverse FFT over N, the central 128 periods. Bar:
series within rel-L2 1e-5 (FP32 transforms against FP64); the uint16 pairs are exactly the
reference's conversion of the returned series, and differ from the FP64 series' codes by at most one
step where the two roundings straddle an integer. End to end, dedispersing the simulated uint16
record with this library's MS-OSM plan at the simulated DM recovers the three-Gaussian profile.

generate_pulsar_signal_new disperses one seed-1 record with one FFT through a causal response: the
inverse FFT of the chirp on filter_point bins (on the device, or in FP64 on the host for designs below
2^12 points and at 2^23), its last Nd taps then tap 0. Its FP64 restatement repeats those steps with
numpy on the same record and chirp; same bar.
"""
import numpy as np
import pytest

import msosm_amd as ms

pytestmark = pytest.mark.gpu


def codes_of(v):
    """The reference's static_cast<uint16_t>(x + 32768.0f) on x86-64, real in the low half."""
    def one(x):
        return (np.trunc(x.astype(np.float32) + np.float32(32768.0)).astype(np.int64) & 0xFFFF).astype(np.uint32)
    return one(v.real) | (one(v.imag) << np.uint32(16))


@pytest.mark.parametrize("case", ["2^22 one workgroup", "2^24 six-step", "2^22 two chunks", "2^26 harness band"])
def test_simulator_matches_fp64_restatement(case):
    bw, dm, f0 = 16e6, 30.0, 1392e6
    if case == "2^26 harness band":  # check_msosm.cpp / check_osm.cpp defaults: 128 MHz, DM 75, 1 GHz, 2.048 ms
        bw, dm, f0 = 128e6, 75.0, 1e9
    period, dummy, valid = {"2^22 one workgroup": (1e-3, 0, 128), "2^24 six-step": (4e-3, 0, 128),
                            "2^22 two chunks": (1e-3, 1, 200), "2^26 harness band": (0.002048, 0, 128)}[case]
    p = ms.simulate_params(bw, dm, f0, period, dummy, valid)
    print(f"case {case}: Np={p.Np} Nd={p.Nd} fftpoint=2^{p.fftpoint.bit_length() - 1} chunks={p.chunks}", flush=True)
    codes, series = ms.simulate_pulsar_block(bw, dm, f0, period, dummy, valid, series=True)
    assert np.array_equal(codes, codes_of(series))
    noise = ms.simulate_noise(bw, dm, f0, period, dummy, valid).astype(np.complex128)
    H = ms.simulate_chirp(bw, dm, f0, p.fftpoint).astype(np.complex128)
    chunk = 128 * p.Np
    for r in range(p.chunks):
        w = noise[r * chunk:r * chunk + p.fftpoint]
        y = np.fft.ifft(np.fft.fft(w) * H)
        n = min(chunk, p.samples - r * chunk)
        ref = y[p.start:p.start + n]
        got = series[r * chunk:r * chunk + n].astype(np.complex128)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        d = np.abs(codes_of(ref).view(np.uint16).astype(np.int32) - codes[r * chunk:r * chunk + n].view(np.uint16).astype(np.int32))
        step = np.minimum(d, 65536 - d)  # codes wrap modulo 2^16
        frac = np.mean(codes_of(ref) != codes[r * chunk:r * chunk + n])
        print(f"  chunk {r}: series rel-L2 {err:.2e}, codes differing {frac:.2e} (max step {step.max()})", flush=True)
        assert err < 1e-5
        assert step.max() <= 1 and frac < 0.05


def new_reference(bw, dm, f0, period, repeat):
    """generate_pulsar_signal_new (simulated_complex_signal.cpp:338-431) in FP64 on the same record."""
    p = ms.simulate_params_new(bw, dm, f0, period, repeat)
    rec = ms.simulate_noise_new(bw, dm, f0, period, repeat).astype(np.complex128)
    h = np.fft.ifft(ms.simulate_chirp(bw, dm, f0, p.filter_point).astype(np.complex128))  # includes 1/filter_point
    tap = np.zeros(p.fftpoint, dtype=np.complex128)
    tap[:p.Nd] = h[p.filter_point - p.Nd:]  # the last Nd taps first, then tap 0
    tap[p.Nd] = h[0]
    y = np.fft.ifft(np.fft.fft(rec) * np.fft.fft(tap))
    return p, y[p.start:p.start + p.samples]


NEW_CASES = {
    # name: (bw, dm, f0, period, repeat); the filter design length decides where it is inverted
    "2^20 record, 2^15 filter on the device": (16e6, 30.0, 1392e6, 1e-3, 32),
    "2^19 record, 2^10 filter on the host": (16e6, 1.0, 1392e6, 1e-3, 16),
    "DM 0, one-tap filter": (16e6, 0.0, 1392e6, 1e-3, 8),
    "2^24 record, 2^23 filter on the host": (128e6, 40.0, 1e9, 0.002048, 4),
    "2^25 record, 2^24 filter six-step": (128e6, 75.0, 1e9, 0.002048, 8),
}


@pytest.mark.parametrize("case", list(NEW_CASES))
def test_simulator_new_matches_fp64_restatement(case):
    bw, dm, f0, period, repeat = NEW_CASES[case]
    p = ms.simulate_params_new(bw, dm, f0, period, repeat)
    print(f"case {case}: Np={p.Np} Nd={p.Nd} fftpoint=2^{p.fftpoint.bit_length() - 1} "
          f"filter_point=2^{p.filter_point.bit_length() - 1}", flush=True)
    codes, series = ms.simulate_pulsar_new(bw, dm, f0, period, repeat, series=True)
    assert np.array_equal(codes, codes_of(series))
    _, ref = new_reference(bw, dm, f0, period, repeat)
    err = np.linalg.norm(series.astype(np.complex128) - ref) / np.linalg.norm(ref)
    d = np.abs(codes_of(ref).view(np.uint16).astype(np.int32) - codes.view(np.uint16).astype(np.int32))
    step = np.minimum(d, 65536 - d)
    frac = np.mean(codes_of(ref) != codes)
    print(f"  series rel-L2 {err:.2e}, codes differing {frac:.2e} (max step {step.max()})", flush=True)
    assert err < 1e-5
    assert step.max() <= 1 and frac < 0.05


@pytest.mark.parametrize("generator", ["block", "new"])
def test_dedispersing_the_simulation_recovers_the_profile(generator):
    # 16 MHz at 1392 MHz, DM 30 (23231 samples of dispersion, 1.45 periods), 1 ms period
    bw, dm, f0, period, nbins = 16e6, 30.0, 1392e6, 1e-3, 250
    if generator == "block":
        codes = ms.simulate_pulsar_block(bw, dm, f0, period, 0, 128)
    else:
        codes = ms.simulate_pulsar_new(bw, dm, f0, period, 128)
    design = ms.Design(bw, f0, [dm])
    batch = 4
    per = batch * design.M
    calls = codes.size // per
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_U16, out_flags=ms.OUT_NONE)
    fold = ms.Fold(period, bw, per, nbins, num_series=1)
    plan.attach_fold(fold)
    for c in range(calls):
        plan.push(np.ascontiguousarray(codes[c * per:(c + 1) * per]))
    prof, sums, cnts = fold.read()
    plan.close()
    fold.close()
    prof = np.asarray(prof, dtype=np.float64).reshape(-1)[:nbins]
    x = (np.arange(nbins) + 0.5) / nbins
    g = (0.45 * np.exp(-0.5 * ((x - 0.455) / 0.008) ** 2) + np.exp(-0.5 * ((x - 0.5) / 0.012) ** 2)
         + 0.28 * np.exp(-0.5 * ((x - 0.555) / 0.006) ** 2)) ** 2
    # the dedispersed series lags the input by one block of M samples: compare over circular shifts
    a = (prof - prof.mean()) / prof.std()
    b = (g - g.mean()) / g.std()
    corr = max(np.dot(np.roll(a, s), b) / nbins for s in range(nbins))
    print(f"folded profile vs three-Gaussian power profile: best correlation {corr:.4f}", flush=True)
    assert corr > 0.97


def pulse_reference(bw, dm, f0, period):
    """generate_pulsar_signal (simulated_complex_signal.cpp:469-611) in closed form, FP64: the chirp on
    `order` bins (w = 2 pi fs i / order stored at (i + order/2) mod order), its inverse FFT placed at
    [order, 2 order), and the circular convolution of impulses of 32767 at r Np + Np/2 and
    r Np + Np/2 + 2 (r < range) with it, which is a sum of shifted copies of the response."""
    p = ms.simulate_params_pulse(bw, dm, f0, period, 1)
    O, Np = p.order, p.Np
    i = np.arange(O, dtype=np.float64)
    step = 2.0 * np.pi * np.float64(np.float32(bw)) / O
    w = step * i
    w0 = 2.0 * np.pi * np.float64(np.float32(f0))
    phase = 4.0 * np.pi * np.pi * 4.15e15 * np.float64(np.float32(dm)) * w * w / ((w + w0) * w0 * w0)
    H1 = np.empty(O, dtype=np.complex64)
    H1[np.where(i < O // 2, O // 2 + i, i - O // 2).astype(np.int64)] = (np.cos(phase) + 1j * np.sin(phase)).astype(np.complex64)
    hpad = np.zeros(2 * O, dtype=np.complex128)
    hpad[O:] = np.fft.ifft(H1.astype(np.complex128))
    j = np.arange(Np)
    out = np.zeros(Np, dtype=np.complex128)
    for r in range(p.range):
        for s in (Np // 2, Np // 2 + 2):
            out += 32767.0 * hpad[(j - r * Np - s) % (2 * O)]
    return p, out


PULSE_CASES = {
    # name: (bw, dm, f0, period); the response length decides where it is designed
    "2^16, response 2^15 in one workgroup": (16e6, 30.0, 1392e6, 1e-3),
    "2^12, response 2^11 on the host": (16e6, 1.0, 1392e6, 1e-4),
    "2^24 six-step, response 2^23 on the host": (128e6, 40.0, 1e9, 0.002048),
    "2^25 six-step, response 2^24 six-step": (128e6, 75.0, 1e9, 0.002048),
}


@pytest.mark.parametrize("case", list(PULSE_CASES))
def test_simulator_pulse_matches_closed_form(case):
    bw, dm, f0, period = PULSE_CASES[case]
    repeat = 3
    p = ms.simulate_params_pulse(bw, dm, f0, period, repeat)
    print(f"case {case}: Np={p.Np} Nd={p.Nd} range={p.range} order=2^{p.order.bit_length() - 1}", flush=True)
    codes, series = ms.simulate_pulsar(bw, dm, f0, period, repeat, series=True)
    assert np.array_equal(codes, codes_of(series))
    per = series[:p.Np]
    for r in range(1, repeat):  # one period, repeated
        assert np.array_equal(series[r * p.Np:(r + 1) * p.Np], per)
    _, ref = pulse_reference(bw, dm, f0, period)
    err = np.linalg.norm(per.astype(np.complex128) - ref) / np.linalg.norm(ref)
    d = np.abs(codes_of(ref).view(np.uint16).astype(np.int32) - codes[:p.Np].view(np.uint16).astype(np.int32))
    step = np.minimum(d, 65536 - d)
    frac = np.mean(codes_of(ref) != codes[:p.Np])
    print(f"  series rel-L2 {err:.2e}, codes differing {frac:.2e} (max step {step.max()})", flush=True)
    assert err < 1e-5
    assert step.max() <= 1 and frac < 0.05
