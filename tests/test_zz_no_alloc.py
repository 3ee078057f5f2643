"""pmt_dot_noalloc (dot products without a workspace; static device scratch) against pmt_dot.

Called through the C-ABI with ctypes.  Two calls in a row, the second with fewer columns and shorter
vectors, so stale partial sums of the first call would show; each result must equal pmt_dot's bits.
Sizes above the maximum column count are refused with PMT_ERR_ARG before anything is launched.
"""
import ctypes

import numpy as np
import pytest
import torch

from pymoto_amd import _lib

pytestmark = pytest.mark.gpu

MAX_COLS = 8  # PMT_DOT_NOALLOC_MAX_COLS


def _ref_dots(lib, X, Y, k, n):
    work = torch.empty(lib.pmt_reduce_work_size(), dtype=torch.float64, device="cuda")
    out = torch.empty(k, dtype=torch.float64, device="cuda")
    for j in range(k):
        rc = lib.pmt_dot(n, ctypes.c_void_p(X.data_ptr() + 8 * j * n), ctypes.c_void_p(Y.data_ptr() + 8 * j * n),
                         ctypes.c_void_p(out.data_ptr() + 8 * j), ctypes.c_void_p(work.data_ptr()), _lib.stream())
        _lib.check(rc, "pmt_dot")
    return out.cpu().numpy()


def _noalloc(lib, X, Y, k, n):
    out = torch.full((k,), np.nan, dtype=torch.float64, device="cuda")
    print(f"pmt_dot_noalloc k={k} n={n}", flush=True)
    rc = lib.pmt_dot_noalloc(k, n, _lib.ptr(X), _lib.ptr(Y), _lib.ptr(out), _lib.stream())
    _lib.check(rc, "pmt_dot_noalloc")
    return out.cpu().numpy()


def test_dot_noalloc_matches_pmt_dot_two_calls(gpu_lib):
    lib = gpu_lib
    rng = np.random.default_rng(4)
    for k, n in ((MAX_COLS, 1_000_003), (3, 4_097)):  # second call smaller: no stale partial may leak in
        X = torch.as_tensor(rng.standard_normal(k * n), device="cuda")
        Y = torch.as_tensor(rng.standard_normal(k * n), device="cuda")
        got = _noalloc(lib, X, Y, k, n)
        want = _ref_dots(lib, X, Y, k, n)
        np.testing.assert_array_equal(got, want)
        # and against the FP64 dot of the same columns on the host (summation order differs: 1e-12)
        Xh, Yh = X.cpu().numpy().reshape(k, n), Y.cpu().numpy().reshape(k, n)
        np.testing.assert_allclose(got, np.einsum("ij,ij->i", Xh, Yh), rtol=1e-10, atol=1e-9)


def test_dot_noalloc_rejects_too_many_columns(gpu_lib):
    lib = gpu_lib
    X = torch.ones(4 * (MAX_COLS + 1), dtype=torch.float64, device="cuda")
    out = torch.zeros(MAX_COLS + 1, dtype=torch.float64, device="cuda")
    rc = lib.pmt_dot_noalloc(MAX_COLS + 1, 4, _lib.ptr(X), _lib.ptr(X), _lib.ptr(out), _lib.stream())
    assert rc == -1 and b"exceed the maximum" in lib.pmt_last_error()
    assert lib.pmt_dot_noalloc(0, 4, _lib.ptr(X), _lib.ptr(X), _lib.ptr(out), _lib.stream()) == -1
    # an empty vector gives zeros
    got = _noalloc(lib, X, X, 2, 0)
    np.testing.assert_array_equal(got, [0.0, 0.0])
