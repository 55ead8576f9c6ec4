"""Writes the golden design tables (design_tables.json) and FFT vectors (fft_vectors.npz) in this folder.

The reference ships no golden vectors, fixtures or known-answer tests for this path. These fixtures
come from an independent restatement, written here in Python, of the reference's design arithmetic
with its float/double mix made explicit (numpy float32 for every C `float` operation, Python floats
for `double`), so that the C oracle and the product's host design are both checked against a second
implementation of the same expressions:

* Nd: `Prepare_MSOSM::Prepare_MSOSM`, dedispersion/msosm/msosm_prepare.cpp:27-33 (members `fs`,
  `kdm`, `dm`, `f0` are float, dedispersion/include/msosm_prepare.h:50-60; `pow` promotes to double);
* minimum order: `calculate_min_order`, msosm_prepare.cpp:57-61 (MIN_SEGMENT_POINTS = 1024,
  msosm_prepare.h:7);
* delaycount and the delay of every bin: `segmentation`, msosm_prepare.cpp:215-284 (w[] is float;
  `pi` is M_PI, utils/include/globals.h:30; `w0 = 2 pi f0` stored as float).

DM 0 gives Nd = 0 and an uninitialised delay table in the reference; it is defined here, as in the
oracle and the product, as delaycount 1 with every delay 0.

FFT vectors: numpy's FFT (an independent library implementation) of seeded random inputs.

Run from anywhere: python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
f32 = np.float32
PI = math.pi
KDM = f32(4.15 * 1e15)  # kdm = 4.15 * 1e15 stored in a float member


def nd(bw, f0, dm):
    bw, f0, dm = f32(bw), f32(f0), f32(dm)
    fs = bw
    prod = f32(KDM * dm)                                   # float * float
    top = f32(f0 + bw)                                     # float + float
    inner = 1.0 / math.pow(float(f0), 2) - 1.0 / math.pow(float(top), 2)  # double
    inv_fs = f32(f32(1) / fs)                              # 1 / fs: int / float -> float
    return int(math.floor(float(prod) * inner / float(inv_fs)))


def min_order(nd_):
    order = math.sqrt(1024 / 2.0) * math.sqrt(nd_)
    return int(math.pow(2, math.ceil(math.log2(order))))


def delaycount(nd_, M):
    if nd_ <= 0:
        return 1
    return int(math.ceil(f32(f32(nd_) / f32(M))))          # ceil(Nd / (float)M) in float


def delays(bw, f0, dm, M):
    """Delay (in blocks) of each of the 2M bins, and the delaycount."""
    n_d = nd(bw, f0, dm)
    dc = delaycount(n_d, M)
    if n_d <= 0:
        return [0] * (2 * M), 1
    bw, f0, dm = f32(bw), f32(f0), f32(dm)
    fs = bw
    w0 = f32(2 * PI * float(f0))
    denom = 4 * PI * PI * float(KDM) * float(dm)
    w = [f32(-1)] + [f32(0)] * (dc - 1) + [f32(1)]
    pos_index = 1
    for j in range(1, dc):
        jm_fs = f32(f32(j * M) / fs)                      # j * M / fs: int / float -> float
        t = 1.0 / math.pow(float(w0), 2) - float(jm_fs) / denom
        w[j] = f32((math.sqrt(1.0 / t) - float(w0) - PI * float(fs)) / PI / float(fs))
        if w[j] < 0:
            pos_index = j + 1
    negcount, poscount = pos_index, dc - pos_index + 1
    segneg = [int(math.floor(f32(f32(w[j] + f32(2)) * f32(M)))) for j in range(negcount)] + [2 * M]
    segpos = [0] + [int(math.floor(f32(w[j + negcount] * f32(M)))) for j in range(poscount)]
    out = [-1] * (2 * M)
    for j in range(negcount):
        for k in range(segneg[j], segneg[j + 1]):
            out[k] = j
    for j in range(poscount):
        for k in range(segpos[j], segpos[j + 1]):
            out[k] = negcount - 1 + j
    return out, dc


def runs(table):
    """Run-length form [[first bin, delay], ...] of a delay table."""
    r = [[0, table[0]]]
    for k in range(1, len(table)):
        if table[k] != table[k - 1]:
            r.append([k, table[k]])
    return r


# (name, bw Hz, band bottom f0 Hz, DM list, FFT length rule): the five workloads' bands and a DM subset
# each, and the reference's own default (benchmark/check_msosm.cpp:12-15)
CASES = [
    ("plain_osm_1400", 16e6, 1392e6, [50.0], "common"),
    ("dm_grid_1250", 64e6, 1218e6, [float(d) for d in np.linspace(0, 100, 512)[::32]] + [57.3, 100.0], "common"),
    ("uneven_800", 32e6, 784e6, [1.0, 10.0, 100.0, 333.0, 500.0, 1000.0], "own"),
    ("stress_350", 50e6, 325e6, [11.71875, 750.0, 1500.0], "common"),
    ("refine_1400", 16e6, 1392e6, [123.456, 122.956, 123.956], "common"),
    ("reference_default", 128e6, 1000e6, [75.0], "common"),
]


def main():
    out = {"generator": "tests/golden/make_golden.py", "cases": []}
    for name, bw, f0, dms, rule in CASES:
        nds = [nd(bw, f0, d) for d in dms]
        ms = [min_order(n) if n > 0 else 0 for n in nds]
        case = {"name": name, "bw": bw, "f0": f0, "dms": dms, "nd": nds, "min_order": ms, "trials": []}
        M_common = max(ms)
        for d, n_d, m_i in zip(dms, nds, ms):
            M = M_common if rule == "common" else max(m_i, 1 << 13)
            table, dc = delays(bw, f0, d, M)
            assert min(table) >= 0, (name, d)
            case["trials"].append({"dm": d, "M": M, "delaycount": dc, "runs": runs(table)})
        out["cases"].append(case)
    with open(os.path.join(HERE, "design_tables.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    rng = np.random.default_rng(20261203)
    vec = {}
    for n in (1, 2, 8, 64, 1024):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        vec[f"x{n}"] = x
        vec[f"fwd{n}"] = np.fft.fft(x)
        vec[f"inv{n}"] = np.fft.ifft(x) * n  # unnormalised e^{+i} sum
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **vec)


if __name__ == "__main__":
    main()
