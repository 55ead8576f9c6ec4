"""Host output pipeline (msosm_saver_*), CPU only: file bytes equal the pushed blocks for io_uring and
pwrite writers, O_DIRECT, several slots; the reference's file naming; I/O failures surface in join
(EACCES at open, short writes and EFBIG under a file-size limit) without hanging the pipeline."""
import os
import resource
import signal
import subprocess
import sys

import numpy as np
import pytest

import msosm_amd as ms


def _expected(pushes, nfiles, block):
    # pushes[k] is [nfiles][block] bytes: file i = concatenation over k of pushes[k][i]
    return [b"".join(p[i].tobytes() for p in pushes) for i in range(nfiles)]


@pytest.mark.parametrize("flags", [ms.SAVE_URING, 0, ms.SAVE_URING | ms.SAVE_DIRECT, ms.SAVE_DIRECT])
@pytest.mark.parametrize("slots", [1, 3])
def test_saved_bytes_equal_pushed_blocks(tmp_path, flags, slots):
    nfiles, block, count = 5, 4096 * 3, 11
    rng = np.random.default_rng(flags * 10 + slots)
    pushes = [rng.integers(0, 256, size=(nfiles, block), dtype=np.uint8) for _ in range(count)]
    paths = [str(tmp_path / f"s{i}.dat") for i in range(nfiles)]
    sv = ms.Saver(paths, block, slots, count, flags)
    sv.start()
    for p in pushes:
        sv.push_host(p)
    sv.join()
    sv.close()
    for i, want in enumerate(_expected(pushes, nfiles, block)):
        with open(paths[i], "rb") as fh:
            assert fh.read() == want


def test_file_names_follow_reference_precision():
    assert ms.saver_file_name("out", "psr", [0.0, 0.5, 1.0], 0) == "out/psr_DM0.0.dat"
    assert ms.saver_file_name("out", "", [10.25, 10.5], 1) == "out/DM10.50.dat"
    assert ms.saver_file_name("", "x", [57.3], 0) == "./x_DM57.3.dat"
    assert ms.saver_file_name("d", "", [12.0, 13.0], 0) == "d/DM12.00.dat"


def test_open_failure_is_reported(tmp_path):
    with pytest.raises(ms.MsosmError) as e:
        ms.Saver([str(tmp_path / "no_such_dir" / "a.dat")], 4096, 2, 1)
    assert e.value.code == ms.ERR_IO and "No such file" in str(e.value)


_CHILD = r"""
import os, sys, signal, resource
sys.path[:0] = sys.argv[1:3]
import numpy as np
import msosm_amd as ms
mode, target = sys.argv[3], sys.argv[4]
ms.lib()  # load the library before any privilege change
if mode == "eacces":
    if os.geteuid() == 0:
        os.setgid(65534); os.setuid(65534)   # drop root so that permissions apply
    try:
        ms.Saver([os.path.join(target, "a.dat")], 4096, 2, 1)
    except ms.MsosmError as e:
        print("ERR", e.code, e, flush=True)
        sys.exit(0)
    print("NOERR", flush=True)
    sys.exit(1)
# size limit: the third block of each file crosses it (short write), the fourth is refused (EFBIG)
signal.signal(signal.SIGXFSZ, signal.SIG_IGN)
block = 4096
resource.setrlimit(resource.RLIMIT_FSIZE, (2 * block + 100, 2 * block + 100))
flags = int(sys.argv[5])
sv = ms.Saver([os.path.join(target, f"f{i}.dat") for i in range(2)], block, 2, 4, flags)
sv.start()
for k in range(4):
    sv.push_host(np.full((2, block), k, dtype=np.uint8))
try:
    sv.join()
except ms.MsosmError as e:
    print("ERR", e.code, e, flush=True)
    sys.exit(0)
print("NOERR", flush=True)
sys.exit(1)
"""


def _child(mode, target, flags=0):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _CHILD, root, os.path.join(root, "acceleratedms-osm_amd"), mode,
                        str(target), str(flags)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout + r.stderr


def test_permission_denied_at_open(tmp_path):
    d = tmp_path / "ro"
    d.mkdir()
    os.chmod(d, 0o555)
    os.chmod(tmp_path, 0o755)
    rc, out = _child("eacces", d)
    assert rc == 0, out
    assert f"ERR {ms.ERR_IO}" in out and "Permission denied" in out


@pytest.mark.parametrize("flags", [ms.SAVE_URING, 0])
def test_short_writes_and_efbig_surface_in_join(tmp_path, flags):
    rc, out = _child("fsize", tmp_path, flags)
    assert rc == 0, out
    assert f"ERR {ms.ERR_IO}" in out and "File too large" in out
    # the bytes below the limit were written in full, including the resumed short write
    for i in range(2):
        with open(tmp_path / f"f{i}.dat", "rb") as fh:
            data = fh.read()
        assert len(data) == 2 * 4096 + 100
        assert data[:4096] == bytes([0]) * 4096 and data[4096:8192] == bytes([1]) * 4096
        assert data[8192:] == bytes([2]) * 100


def test_read_complex_bin_decodes_saved_words(tmp_path):
    # python_check/check.py:5-53: low 16 bits real, high 16 imaginary, code - 32768, code 0 read as 0
    codes = np.array([0x0000, 0x8000, 0x8001, 0x7FFF, 0xFFFF, 0x0001], dtype=np.uint32)
    values = np.array([0.0, 0.0, 1.0, -1.0, 32767.0, -32767.0])
    words = (codes[:, None] | (codes[None, ::-1] << np.uint32(16))).reshape(-1)  # every real/imag pairing
    expect = (values[:, None] + 1j * values[None, ::-1]).reshape(-1)
    nblock = 1024  # one saver block of uint32 words, written through the pwrite writer
    block = np.zeros(nblock, dtype=np.uint32)
    block[:words.size] = words
    path = str(tmp_path / "x_DM1.dat")
    sv = ms.Saver([path], block.nbytes, 1, 1, 0)
    sv.start()
    sv.push_host(block.view(np.uint8).reshape(1, -1))
    sv.join()
    sv.close()
    got = ms.read_complex_bin(path)
    assert got.dtype == np.complex128 and got.size == nblock
    assert np.array_equal(got[:words.size], expect) and not np.any(got[words.size:])
    # start_index and length count samples (4 bytes each)
    assert np.array_equal(ms.read_complex_bin(path, 7, 5), expect[7:12])
    assert np.array_equal(ms.read_complex_bin(path, 30), np.concatenate([expect[30:], np.zeros(nblock - 36)]))


@pytest.mark.parametrize("uring", [True, False])
def test_caller_threads_run_the_poll_and_write_loops(tmp_path, uring):
    # poll_cuda_state / save_to_disk / save_to_disk_pwrite (msosm_data_save_uring.cpp:115-258) run on the
    # caller's threads instead of start_saving's: one poller, one writer per slot, the producer on the
    # main thread; the files hold every push's block of each series at its offset
    import threading

    nfiles, block, slots, pushes = 2, 4096, 3, 7
    paths = [str(tmp_path / f"s{i}.dat") for i in range(nfiles)]
    sv = ms.Saver(paths, block, slots, pushes, flags=ms.SAVE_URING if uring else 0)
    errors = []

    def run(fn, *a):
        try:
            fn(*a)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(sv.run_poller,))]
    threads += [threading.Thread(target=run, args=(sv.run_writer, k, uring)) for k in range(slots)]
    for t in threads:
        t.start()
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, size=(pushes, nfiles, block), dtype=np.uint8)
    for k in range(pushes):
        sv.push_host(data[k])
    for t in threads:
        t.join(timeout=60)
        assert not t.is_alive()
    assert not errors, errors
    sv.join()
    sv.close()
    for i, p in enumerate(paths):
        got = np.fromfile(p, dtype=np.uint8)
        assert np.array_equal(got, data[:, i, :].reshape(-1))


def test_caller_loops_refuse_misuse(tmp_path):
    sv = ms.Saver([str(tmp_path / "a.dat")], 4096, 2, 2, flags=0)
    with pytest.raises(ms.MsosmError, match="MSOSM_SAVE_URING"):
        sv.run_writer(0, True)
    with pytest.raises(ms.MsosmError, match="out of range"):
        sv.run_writer(5, False)
    sv.close()
    sv = ms.Saver([str(tmp_path / "b.dat")], 4096, 2, 2)
    sv.start()
    with pytest.raises(ms.MsosmError, match="msosm_saver_start runs the loops"):
        sv.run_poller()
    for k in range(2):
        sv.push_host(np.zeros(4096, np.uint8))
    sv.join()
    sv.close()
