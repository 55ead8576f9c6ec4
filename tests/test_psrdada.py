"""DADA reader of the reference-compatible API (PSRDADA, psrdada/psrdada.cpp), CPU only: header
parsing, readCount and the per-polarisation frame de-interleave, through the C++ driver."""
import os
import subprocess

import numpy as np
import pytest

from tests.util import write_dada

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "check_reference_api")


def test_header_and_frame_deinterleave(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (make -C tests/cpp)")
    n = 4096 * 3  # three reads of 4096 samples per polarisation
    k = np.arange(n, dtype=np.uint32)
    pol1 = ((k + 1) & 0xFFFF) | (((k * 3 + 7) & 0xFFFF) << 16)
    pol2 = ((k * 5 + 2) & 0xFFFF) | (((k + 9) & 0xFFFF) << 16)
    path = tmp_path / "obs_01.dada"
    write_dada(path, pol1, pol2, 16.0, 1400.0)
    r = subprocess.run([DRIVER, "dada_info", str(path), "4096"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    head = lines[0].split()
    vals = dict(zip(head[0::2], head[1::2]))
    assert float(vals["bw"]) == 16.0 and float(vals["fc"]) == 1400.0
    assert vals["nbit"] == "16" and vals["nchan"] == "1" and vals["ndim"] == "2" and vals["npol"] == "2"
    assert vals["readCount"] == "3"
    assert vals["out"] == str(tmp_path / "obs_01_dedisp_folded.txt")
    for i in range(4):
        s = [int(v) for v in lines[1 + i].split()[1:]]
        assert s == [pol1[i] & 0xFFFF, pol1[i] >> 16, pol2[i] & 0xFFFF, pol2[i] >> 16]
    last = [int(v) for v in lines[5].split()[1:]]
    assert last == [pol1[2048] & 0xFFFF, pol2[2048] & 0xFFFF]  # second frame of each polarisation
