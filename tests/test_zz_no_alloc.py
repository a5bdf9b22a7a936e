"""The allocation-free dot (qblas_hip_dot_noalloc) through the C ABI against qblas_hip_dot.

For each mode it has (REFERENCE, EXACT, DD), two calls in a row on the same GPU, the second with a
smaller input, each bit-identical to qblas_hip_dot on the same data (which is itself checked
against the oracle elsewhere); the EXACT form also on data whose products span 2^+-12000 (the
full-range path, whose records live in the static scratch too).  A call that needs more scratch
than the library holds (REFERENCE with T = 100000: 3.2 MB) is refused with QBLAS_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, to_dev

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _noalloc(Q):
    L = Q.lib()
    fn = L.qblas_hip_dot_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return fn


def _run(Q, cuda, fn, mode, n, x, incx, y, incy, T):
    dx, dy = to_dev(x), to_dev(y)
    r_ref, r_na = dev_empty(1), dev_empty(1)
    Q.dev_dot(mode, n, dx.data_ptr(), incx, dy.data_ptr(), incy, T, r_ref.data_ptr())
    rc = fn(mode, n, dx.data_ptr(), incx, dy.data_ptr(), incy, T, r_na.data_ptr(), None)
    assert rc == 0, Q.last_error()
    cuda.cuda.synchronize()
    return pair(from_dev(r_na)[0]), pair(from_dev(r_ref)[0])


@pytest.mark.parametrize("mode_name", ["reference", "exact", "dd"])
def test_noalloc_dot_matches_dot(O, Q, cuda, mode_name):
    fn = _noalloc(Q)
    mode = Q.MODES[mode_name]
    for n, incx, incy, T, seed in ((1 << 22, 1, 1, 16, 3), (5001, 2, 3, 7, 4)):
        _say(f"noalloc dot mode={mode_name} n={n} incx={incx} incy={incy} T={T}")
        x = O.fill(0, seed, (n - 1) * incx + 1)
        y = O.fill(0, seed + 1, (n - 1) * incy + 1)
        got, want = _run(Q, cuda, fn, mode, n, x, incx, y, incy, T)
        assert got == want


def test_noalloc_exact_full_range_and_limits(O, Q, cuda):
    fn = _noalloc(Q)
    for n, seed in ((1 << 20, 11), (777, 12)):
        _say(f"noalloc exact dot, products over 2^+-12000, n={n}")
        x = O.fill(1, seed, n, -6000, 6000)
        y = O.fill(1, seed + 1, n, -6000, 6000)
        got, want = _run(Q, cuda, fn, Q.MODE_EXACT, n, x, 1, y, 1, 1)
        assert got == want
        assert got == O.exact_dot(x, y)[0]
    x = O.fill(0, 1, 1000)
    dx, dr = to_dev(x), dev_empty(1)
    assert fn(Q.MODE_REFERENCE, 1000, dx.data_ptr(), 1, dx.data_ptr(), 1, 100000, dr.data_ptr(), None) == 1
    assert "scratch" in Q.last_error()
    Q.lib().qblas_clear_error()
    assert fn(Q.MODE_OZAKI, 1000, dx.data_ptr(), 1, dx.data_ptr(), 1, 1, dr.data_ptr(), None) == 3
    Q.lib().qblas_clear_error()
    # n = 0 gives +0 (level1.hpp:83-84)
    assert fn(Q.MODE_EXACT, 0, dx.data_ptr(), 1, dx.data_ptr(), 1, 1, dr.data_ptr(), None) == 0
    cuda.cuda.synchronize()
    assert pair(from_dev(dr)[0]) == (0, 0)
