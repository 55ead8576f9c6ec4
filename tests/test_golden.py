"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py) against the C oracle and the
product's host design.

The design tables come from an independent Python restatement of the reference's float/double design
arithmetic (msosm_prepare.cpp:27-33, 57-61, 215-284); the oracle (C) and the library (C++) must
reproduce every Nd, minimum order, delaycount and per-bin delay bit for bit. The FFT vectors are
numpy's transforms of seeded inputs; the oracle's FP64 FFT must match them to rounding.
"""
import json
import os

import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "design_tables.json")) as _f:
    TABLES = json.load(_f)
CASES = {c["name"]: c for c in TABLES["cases"]}


def expand(runs, L):
    out = np.empty(L, dtype=np.int64)
    for (start, delay), nxt in zip(runs, runs[1:] + [[L, None]]):
        out[start:nxt[0]] = delay
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_design_matches_golden(name):
    c = CASES[name]
    for dm, n_d, m_i, tr in zip(c["dms"], c["nd"], c["min_order"], c["trials"]):
        assert orc.nd(c["bw"], c["f0"], dm) == n_d
        if n_d > 0:
            assert orc.min_order(n_d) == m_i
        assert orc.delaycount(n_d, tr["M"]) == tr["delaycount"]
        delays, dc = orc.segmentation(c["bw"], c["f0"], float(np.float32(dm)), tr["M"])
        assert dc == tr["delaycount"]
        assert np.array_equal(delays, expand(tr["runs"], 2 * tr["M"])), (name, dm)


@pytest.mark.parametrize("name", list(CASES))
def test_library_design_matches_golden(name):
    c = CASES[name]
    for dm, n_d, m_i, tr in zip(c["dms"], c["nd"], c["min_order"], c["trials"]):
        assert ms.design_nd(c["bw"], c["f0"], dm) == n_d
        if n_d > 0:
            assert ms.design_min_order(n_d) == m_i
        assert ms.design_delaycount(n_d, tr["M"]) == tr["delaycount"]
        delays, dc = ms.design_segmentation(c["bw"], c["f0"], dm, tr["M"])
        assert dc == tr["delaycount"]
        assert np.array_equal(np.asarray(delays, dtype=np.int64), expand(tr["runs"], 2 * tr["M"])), (name, dm)


def test_golden_tables_satisfy_the_segmentation_invariants():
    # every bin has a delay; delay 0 at the band bottom (bin M), delaycount-1 at the top (bin M-1),
    # non-decreasing along M..2M-1, 0..M-1 in unit steps
    for c in TABLES["cases"]:
        for tr in c["trials"]:
            M = tr["M"]
            d = expand(tr["runs"], 2 * M)
            walk = np.concatenate([d[M:], d[:M]])
            assert walk[0] == 0 and walk[-1] == tr["delaycount"] - 1
            assert set(np.diff(walk)) <= {0, 1}


@pytest.mark.parametrize("n", [1, 2, 8, 64, 1024])
def test_oracle_fft_matches_numpy_vectors(n):
    v = np.load(os.path.join(GOLDEN, "fft_vectors.npz"))
    x = v[f"x{n}"]
    scale = max(1.0, np.abs(v[f"fwd{n}"]).max())
    assert np.abs(orc.fft(x, -1) - v[f"fwd{n}"]).max() < 1e-12 * scale * max(1, np.log2(n))
    assert np.abs(orc.fft(x, +1) - v[f"inv{n}"]).max() < 1e-12 * scale * max(1, np.log2(n))
