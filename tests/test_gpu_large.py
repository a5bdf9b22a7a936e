"""GPU launches over more than 2^32 work-items.

A launch dimension holds at most 2^32 - 1 work-items, and a larger one silently covers only the
count modulo 2^32. Element-wise kernels therefore stride over their elements from a capped grid,
and the one-wave-per-row gemv strides over its rows.
  - The seeded fill of 2^32 + 4096 entries (69 GB) matches the oracle's sequence at its start,
    across index 2^32, and at its end.
  - A RowMajor EXACT gemv with 2^26 + 96 rows of 64 columns (69 GB, one wave per row) is
    correctly rounded on its first and last rows.
"""
import numpy as np
import pytest

from helpers import from_dev

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_fill_beyond_2_32_entries(O, Q, cuda):
    n = (1 << 32) + 4096
    _say(f"fill of {n} entries")
    buf = cuda.full((n, 2), -1, dtype=cuda.int64, device="cuda:0")
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 9, n, buf.data_ptr())
    cuda.cuda.synchronize()
    for start, cnt in ((0, 256), ((1 << 32) - 128, 256), (n - 4096, 4096)):
        got = from_dev(buf[start:start + cnt])
        assert np.array_equal(got, O.fill(0, 9, cnt, start=start)), start
    del buf
    cuda.cuda.empty_cache()


def test_exact_gemv_tall_beyond_2_32_lanes(O, Q, cuda):
    m, n = (1 << 26) + 96, 64
    _say(f"exact gemv RowMajor {m} x {n}")
    A = cuda.empty((m * n, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 21, m * n, A.data_ptr())
    x = O.fill(0, 22, n)
    y0 = O.fill(0, 23, m)
    dx = cuda.from_numpy(x.view(np.int64)).to("cuda:0")
    dy = cuda.from_numpy(y0.view(np.int64)).to("cuda:0")
    Q.dev_gemv(Q.MODE_EXACT, 0, m, n, 1.5, A.data_ptr(), n, dx.data_ptr(), 1, -0.5, dy.data_ptr(), 1)
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() == 0
    for r0, rows in ((0, 64), (m - 160, 160)):
        Ar = O.fill(0, 21, rows * n, start=r0 * n)
        want = y0[r0:r0 + rows].copy()
        O.exact_gemv(0, rows, n, 1.5, Ar, n, x, 1, -0.5, want, 1)
        got = from_dev(dy[r0:r0 + rows])
        assert np.array_equal(got, want), r0
    del A
    cuda.cuda.empty_cache()
