"""Pins for the DM-derivative pieces of the oracle and the library's dH/dDM design (no GPU).

* dH/dDM against central differences of the FP64 design in DM, at two step sizes: the error must
  fall by 4x when h halves (second-order differences), and the Richardson combination must agree
  to < 1e-5. DM and DM +- h are exact floats, so the float DM argument adds nothing.
* The folded-profile derivative (MS-OSM with dH/dDM, then 2 Re(conj(y) dy) folded) against central
  differences of the oracle's whole profile pipeline with the segmentation frozen at the start DM.
* The library's complex64 dH/dDM against the oracle's, to complex64 rounding.
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc

BW, F0 = 16e6, 1392e6


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("dm", [0.0, 37.5, 100.0])
def test_ddm_design_matches_central_differences(dm):
    M = 2048
    T = lambda d: orc.transfer_function(BW, F0, d, M)
    dH = orc.transfer_function_ddm(BW, F0, dm, M)
    fd = lambda h: (T(dm + h) - T(dm - h)) / (2 * h)
    a, b = fd(2.0 ** -14), fd(2.0 ** -15)
    ea, eb = _rel(a, dH), _rel(b, dH)
    assert 3.5 < ea / eb < 4.5, (ea, eb)
    assert _rel((4 * b - a) / 3, dH) < 1e-5


def test_ddm_is_not_the_design_at_another_dm():
    # a plausible mistake (dropping -i, or the factor phi/DM) fails the central-difference pin; check
    # the sign convention directly: dc/dDM = -i phi_1 c, so dH at DM 0 is -i times the DM-1 phase map
    M = 1024
    dH = orc.transfer_function_ddm(BW, F0, 0.0, M)
    # even bins: H[2n] = c_n exactly, so dH[2n] = dc_n/dDM = -i phi_1(n) at DM 0
    n = np.arange(M)
    j = np.where(n < M // 2, n + M // 2, n - M // 2)
    w = 2 * np.pi * BW / M * j
    w0 = 2 * np.pi * np.float64(np.float32(F0))
    phi1 = 4 * np.pi ** 2 * 4.15e15 * w * w / ((w + w0) * w0 * w0)
    assert np.allclose(dH[0::2], -1j * phi1, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("dm", [0.0, 57.3])
def test_library_ddm_design_rounding(dm):
    M = 4096
    ref = orc.transfer_function_ddm(BW, F0, float(np.float32(dm)), M)
    got = ms.design_transfer_function_ddm(BW, F0, dm, M).astype(np.complex128)
    assert np.max(np.abs(got - ref)) <= 2e-7 * np.max(np.abs(ref))


def test_profile_tangent_matches_central_differences():
    # a small MS-OSM (M=1024, 3 delay blocks) of a dispersed pulse, DM perturbed around 10
    dm0, M, nbins, period = 10.0, 1024, 64, 0.0011
    n = 1 << 15
    x = orc.pulsar(n, BW, F0, dm0, period, duty=0.05, amp=0.8, sigma=1.0, seed=21)
    delays, dc = orc.segmentation(BW, F0, dm0, M)
    assert dc > 1
    dt = 1.0 / BW

    def profile(dm):
        y = orc.msosm(x, M, delays[None], orc.transfer_function(BW, F0, dm, M)[None])[0]
        p, _, _ = orc.profile_tangent(y, np.zeros_like(y), nbins, 0.0, dt, period)
        return p

    y = orc.msosm(x, M, delays[None], orc.transfer_function(BW, F0, dm0, M)[None])[0]
    dy = orc.msosm(x, M, delays[None], orc.transfer_function_ddm(BW, F0, dm0, M)[None])[0]
    p, dp, counts = orc.profile_tangent(y, dy, nbins, 0.0, dt, period)
    assert counts.sum() == n
    fd = lambda h: (profile(dm0 + h) - profile(dm0 - h)) / (2 * h)
    a, b = fd(2.0 ** -14), fd(2.0 ** -15)
    rich = (4 * b - a) / 3
    assert 3.5 < _rel(a, dp) / _rel(b, dp) < 4.5
    assert _rel(rich, dp) < 5e-5, (_rel(a, dp), _rel(b, dp), _rel(rich, dp))
    # linearity of MS-OSM in H is what makes dy the derivative series: check it directly
    H2 = orc.transfer_function(BW, F0, dm0 + 1.0, M)
    y2 = orc.msosm(x, M, delays[None], H2[None])[0]
    ysum = orc.msosm(x, M, delays[None], (orc.transfer_function(BW, F0, dm0, M) + 2.0 * H2)[None])[0]
    assert _rel(ysum, y + 2.0 * y2) < 1e-13
