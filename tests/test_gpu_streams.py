"""Save-mode pipeline on the plan's copy streams (msosm_gpu_batch.cpp:15-27,203-245): host input copied
asynchronously on the copy-in stream, uint16 output (two device buffers) copied out after every push on
the copy-out stream. The output bytes are identical to the serialised form (synchronous input copy, each
push's output read back before the next push)."""
import ctypes as C

import numpy as np
import pytest

import msosm_amd as ms
from tests.util import make_input

pytestmark = pytest.mark.gpu

BW, F0 = 64e6, 1218e6


def _pinned(nbytes):
    p = C.c_void_p()
    ms._check(ms.lib().msosm_host_alloc(C.c_size_t(nbytes), C.byref(p)))
    return p


def test_save_mode_streams_match_serialised_pushes():
    dms = np.linspace(10.0, 100.0, 24).astype(np.float32)
    design = ms.Design(BW, F0, dms)
    batch, calls = 4, 12
    per = batch * design.M
    n = per * calls
    raw, _ = make_input(ms.IN_I8, 1 << (n - 1).bit_length(), BW, F0, 100.0, 0.0011, seed=81)
    raw = np.ascontiguousarray(raw[:n])
    out_bytes = len(dms) * per * 4
    print(f"case save-mode streams D={len(dms)} M={design.M} batch={batch} pushes={calls}", flush=True)
    # serialised: synchronous input copy, output read back (and synchronised) after each push
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_U16)
    ref = []
    for c in range(calls):
        plan.push(raw[c * per:(c + 1) * per])
        ref.append(plan.output_u16())
    plan.close()
    # pipelined: pinned input pushed asynchronously, every push's output copied into its own pinned slot,
    # one synchronize at the end
    lib = ms.lib()
    pin_in = _pinned(raw.nbytes)
    C.memmove(pin_in, raw.ctypes.data, raw.nbytes)
    slots = [_pinned(out_bytes) for _ in range(calls)]
    plan = ms.Plan.from_design(design, batch, in_format=ms.IN_I8, out_flags=ms.OUT_U16)
    for c in range(calls):
        ms._check(lib.msosm_plan_push(plan.h, C.c_void_p(pin_in.value + c * per * 2), ms.INPUT_HOST_ASYNC))
        ms._check(lib.msosm_plan_get_output(plan.h, ms.OUT_U16, slots[c]))
    plan.synchronize()
    plan.close()
    for c in range(calls):
        got = np.ctypeslib.as_array(C.cast(slots[c], C.POINTER(C.c_uint32)), shape=(out_bytes // 4,))
        assert np.array_equal(got, ref[c].ravel()), f"push {c}: pipelined output differs"
    for p in [pin_in] + slots:
        lib.msosm_host_free(p)
