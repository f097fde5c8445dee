"""The allocation-free row log-softmax (ac_row_log_softmax_static) against the allocating one, through the C ABI.

ac_row_log_softmax_static runs the same three kernels as ac_row_log_softmax, with its partials in a fixed
device-resident scratch instead of a stream-ordered allocation, and the same chunking; so its output must equal the
allocating entry point's bit for bit.  Two calls in a row, the second on a smaller input (the scratch keeps the first
call's contents, which must not leak into the second), then the size limit.
"""
import ctypes
import sys

import numpy as np
import pytest

from application_classification_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(fn, x):
    d = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    rc = fn(x.shape[0], x.shape[1], ctypes.c_void_p(d.data_ptr()), None)
    torch.cuda.synchronize()
    assert rc == _lib.AC_OK, _lib.lib().ac_global_error()
    return d.cpu().numpy()


def test_static_scratch_row_log_softmax_matches_allocating():
    L = _lib.lib()
    r = np.random.default_rng(7)
    for n_row, n_col in ((24, 1_500_003), (3, 40_000)):
        x = (r.normal(size=(n_row, n_col)) * 20).astype(np.float32)
        x[n_row - 1, 5] = 150.0  # a dominant entry: its log-probability is tiny
        say(f"static scratch: {n_row} x {n_col}")
        got = run(L.ac_row_log_softmax_static, x)
        want = run(L.ac_row_log_softmax, x)
        np.testing.assert_array_equal(got, want)
        assert np.abs(np.log(np.exp(got.astype(np.float64)).sum(axis=1))).max() < 1e-5
    big = np.zeros((40_000, 2), np.float32)
    d = torch.from_numpy(big).cuda()
    rc = L.ac_row_log_softmax_static(big.shape[0], big.shape[1], ctypes.c_void_p(d.data_ptr()), None)
    assert rc == _lib.AC_ERR_ARG and b"rows" in L.ac_global_error()
