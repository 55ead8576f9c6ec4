"""The fold entry point that allocates nothing (msosm_fold_series_noalloc) against the fold object's
entry points (msosm_fold_create + msosm_fold_accumulate + msosm_fold_read) on the same device series.

Both run the same fold kernels; the sums are FP64 atomics whose order may differ between launches, so
the bar is: counts bit-exact, profiles within 1e-6 relative (float rounding of an FP64 ratio whose last
bits may differ). Two calls in a row on one GPU share the static scratch array; the second, smaller one
must not see the first one's partials.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import msosm_amd as ms

pytestmark = pytest.mark.gpu


def _object_fold(x, size, stride, nseries, nbins, fs, period):
    """A fresh fold object's first call (its clock starts at t = 0)."""
    fold = ms.Fold(period, fs, size, nbins, num_series=nseries, device=0)
    lib = ms.lib()
    ms._check(lib.msosm_fold_accumulate(fold.h, C.c_void_p(x.data_ptr()), None, C.c_longlong(stride), None))
    prof, sums, cnts = fold.read()
    fold.close()
    return prof, cnts


def _noalloc(x, size, stride, nseries, nbins, t_start, fs, period):
    prof = torch.full((nseries, nbins), -1.0, dtype=torch.float32, device="cuda:0")
    cnts = torch.full((nseries, nbins), -1, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    lib = ms.lib()
    ms._check(lib.msosm_fold_series_noalloc(C.c_void_p(x.data_ptr()), None, C.c_longlong(size),
                                            C.c_longlong(stride), nseries, C.c_ulong(nbins), C.c_double(t_start),
                                            C.c_double(fs), C.c_double(period), C.c_void_p(prof.data_ptr()),
                                            C.c_void_p(cnts.data_ptr()), None))
    torch.cuda.synchronize()
    return prof.cpu().numpy(), cnts.cpu().numpy().view(np.uint64)


def test_noalloc_fold_matches_fold_object_on_two_calls():
    fs, period = 64e6, 5.0e-3
    g = torch.Generator(device="cuda:0").manual_seed(99)
    cases = [(512, 1 << 18, 1024), (64, 1 << 16, 256)]  # (series, samples, bins): the second is smaller
    for nseries, size, nbins in cases:
        stride = size + 64
        x = torch.randn(nseries * stride, dtype=torch.complex64, device="cuda:0", generator=g)
        x[:size // 7] = 0  # zero-intensity samples are skipped, as in Fold_GPU
        print(f"case no-alloc fold series={nseries} size={size} nbins={nbins}", flush=True)
        p0, c0 = _object_fold(x, size, stride, nseries, nbins, fs, period)
        p1, c1 = _noalloc(x, size, stride, nseries, nbins, 0.0, fs, period)
        assert np.array_equal(c0, c1)
        np.testing.assert_allclose(p1, p0, rtol=1e-6, atol=0)
        print(f"  counts equal, max profile rel diff {np.max(np.abs(p1 - p0) / np.maximum(np.abs(p0), 1e-30)):.2e}",
              flush=True)


def test_noalloc_fold_rejects_sizes_beyond_its_scratch():
    lib = ms.lib()
    x = torch.zeros(16, dtype=torch.complex64, device="cuda:0")
    out = torch.zeros(16, dtype=torch.float32, device="cuda:0")
    rc = lib.msosm_fold_series_noalloc(C.c_void_p(x.data_ptr()), None, C.c_longlong(16), C.c_longlong(16), 2048,
                                       C.c_ulong(1024), C.c_double(0.0), C.c_double(1e6), C.c_double(1e-3),
                                       C.c_void_p(out.data_ptr()), None, None)
    assert rc == ms.ERR_ARG
    assert b"exceeds the static scratch" in lib.msosm_last_error()
