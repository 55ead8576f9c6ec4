"""The reference's correctness protocol (benchmark/check_osm.cpp:73-154, benchmark/check_msosm.cpp:89-123,
182-219): dedisperse a simulated pulsar with plain OSM (complex64 output) and with MS-OSM (save mode,
uint16-pair output), skip the common warm-up (2 nextpow2(Nd) samples, the OSM order, rounded up to a
pulse period), fold |x| over the same simulated periods at one bin per sample, and compare the two
profiles.

The record comes from the library's GPU form of generate_pulsar_signal_block (the harnesses' generator,
simulated_complex_signal.cpp:74-277); both methods read the same samples. Each GPU profile agrees with
the FP64 oracle's profile of the same method (parity, rel-L2 <= 1e-5 for OSM; the uint16 codes add their
rounding for MS-OSM), and the two methods agree with each other as closely as they do in the oracle: the
OSM/MS-OSM difference is a property of the algorithms, not of the GPU.
"""
import numpy as np
import pytest

import msosm_amd as ms
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def fold_abs(x, fold_start, block, repeat):
    """check_osm.cpp:139-151 / check_msosm.cpp:182-200: mean |x| per sample phase over `repeat` periods."""
    seg = x[fold_start:fold_start + repeat * block].reshape(repeat, block)
    return np.mean(np.abs(seg.astype(np.complex128)), axis=0)


def test_osm_and_msosm_profiles_agree_over_the_common_periods():
    # a 1.024 ms period: 16384 samples, so the generator's chunk transform is 2^22 points
    bw, f0, dm, period = 16e6, 1392e6, 50.0, 0.001024
    block = int(period * bw)
    nd = ms.design_nd(bw, f0, dm)
    m_osm = 2 * (1 << (nd - 1).bit_length())         # Prepare_OSM: 2 nextpow2(Nd) (osm_prepare.cpp:17-30)
    osm = ms.Design(bw, f0, [dm], M=m_osm)
    msosm = ms.Design(bw, f0, [dm])
    assert osm.delaycount == 1 and msosm.M < m_osm
    b_osm, b_ms = 2, 32
    repeat = 128
    common = m_osm                                   # check_msosm.cpp:101-110
    fold_start = -(-common // block) * block
    need = fold_start + repeat * block
    per_osm, per_ms = b_osm * m_osm, b_ms * msosm.M
    n = max(-(-need // per_osm) * per_osm, -(-need // per_ms) * per_ms)
    periods = -(-n // block)
    print(f"case protocol Nd={nd} OSM L={2 * m_osm} MS-OSM L={2 * msosm.M} dc={msosm.delaycount} "
          f"fold_start={fold_start} repeat={repeat} periods={periods}", flush=True)
    codes = ms.simulate_pulsar_block(bw, dm, f0, period, periods - repeat, repeat)
    assert codes.size >= n
    codes = codes[:n]
    x = orc.unpack_u16(codes)
    # OSM, complex64 output
    plan = ms.Plan.from_design(osm, b_osm, in_format=ms.IN_U16, out_flags=ms.OUT_C64)
    y_osm = []
    for c in range(n // per_osm):
        plan.push(codes[c * per_osm:(c + 1) * per_osm])
        y_osm.append(plan.output_c64()[0])
    plan.close()
    y_osm = np.concatenate(y_osm)
    # MS-OSM, uint16-pair output (the harness's save mode)
    plan = ms.Plan.from_design(msosm, b_ms, in_format=ms.IN_U16, out_flags=ms.OUT_U16)
    y_ms = []
    for c in range(n // per_ms):
        plan.push(codes[c * per_ms:(c + 1) * per_ms])
        y_ms.append(plan.output_u16()[0])
    plan.close()
    y_ms = orc.unpack_u16(np.concatenate(y_ms))
    p_osm = fold_abs(y_osm, fold_start, block, repeat)
    p_ms = fold_abs(y_ms, fold_start, block, repeat)
    # the same protocol on the FP64 oracle's series
    r_osm = orc.msosm(x, m_osm, osm.delays, osm.H.astype(np.complex128))[0]
    r_ms = orc.msosm(x, msosm.M, msosm.delays, msosm.H.astype(np.complex128))[0]
    q_osm = fold_abs(r_osm, fold_start, block, repeat)
    q_ms = fold_abs(r_ms, fold_start, block, repeat)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    e_osm, e_ms = rel(p_osm, q_osm), rel(p_ms, q_ms)
    d_gpu, d_orc = rel(p_ms, p_osm), rel(q_ms, q_osm)
    corr = float(np.corrcoef(p_ms, p_osm)[0, 1])
    peak = (int(np.argmax(p_osm)), int(np.argmax(p_ms)))
    print(f"  OSM profile vs oracle {e_osm:.2e}; MS-OSM (uint16) profile vs oracle {e_ms:.2e}", flush=True)
    print(f"  MS-OSM vs OSM: GPU {d_gpu:.3e}, oracle {d_orc:.3e}; correlation {corr:.6f}; peaks {peak}", flush=True)
    assert e_osm < 1e-5
    assert e_ms < 2e-4        # one uint16 code step is ~1e-4 of the output rms here
    assert abs(d_gpu - d_orc) < 2e-4 + 0.05 * d_orc
    assert corr > 0.99 and abs(peak[0] - peak[1]) <= 2
