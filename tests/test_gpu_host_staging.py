"""Host-pointer entry points over the pipelined staging of caller buffers.

A pageable host span of 2 MiB or more is copied in 8 MiB chunks by up to 8 threads through pinned
staging slots; a pinned one (hipHostMalloc, torch's pin_memory) by one DMA.  Each case runs the
host-pointer entry point and the device-pointer one (or the oracle) on the same data and requires
the same bits, including the entries between the rows of C (ldc > n) and between strided elements
of y, which the host call writes back unchanged:
  - spans of 2.3 to 22 MiB: one to three chunks over as many workers, the last chunk short;
  - outputs (C, y) large enough to be staged on the way back too;
  - pageable and pinned caller buffers;
  - a second call right after the first, reusing the staging slots.
"""
import ctypes

import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, to_dev

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _host(arr, pinned, torch):
    """A host copy of arr: pageable (numpy) or pinned (torch pin_memory); returns (array, owner)."""
    if not pinned:
        return arr.copy(), None
    t = torch.empty(arr.shape, dtype=torch.int64, pin_memory=True)
    h = t.numpy().view(np.uint64)
    h[...] = arr
    return h, t


class _mode:
    def __init__(self, Q, mode):
        self.Q, self.mode = Q, mode

    def __enter__(self):
        self.saved = self.Q.get_mode()
        self.Q.set_mode(self.mode)

    def __exit__(self, *exc):
        self.Q.set_mode(self.saved)


@pytest.mark.parametrize("mode", ["dd", "ozaki"])
@pytest.mark.parametrize("pinned", [False, True])
def test_host_gemm_staged(O, Q, cuda, mode, pinned):
    m, n, k = 700, 1100, 900
    lda, ldb, ldc = k + 5, n + 2, n + 3
    _say(f"host gemm {mode} pinned={pinned} {m}x{n}x{k}")
    A = O.fill(0, 31, (m - 1) * lda + k)
    B = O.fill(0, 32, (k - 1) * ldb + n)
    C0 = O.fill(0, 33, (m - 1) * ldc + n)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C0)
    Q.dev_gemm(Q.MODES[mode], 0, m, n, k, 1.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, -0.5, dC.data_ptr(), ldc)
    cuda.cuda.synchronize()
    want = from_dev(dC)
    (hA, oA), (hB, oB), (hC, oC) = (_host(a, pinned, cuda) for a in (A, B, C0))
    with _mode(Q, mode):
        for _ in range(2):
            hC[...] = C0
            Q.gemm(0, m, n, k, 1.5, hA, lda, hB, ldb, -0.5, hC, ldc)
            assert np.array_equal(hC, want)


@pytest.mark.parametrize("mode", ["reference", "dd"])
@pytest.mark.parametrize("pinned", [False, True])
def test_host_gemv_staged(O, Q, cuda, mode, pinned):
    # ColMajor, tall: A (22 MiB) and y with incy = 2 (6.1 MiB) are staged both ways
    m, n, incy = 200003, 7, 2
    _say(f"host gemv {mode} pinned={pinned} ColMajor {m}x{n} incy={incy}")
    A = O.fill(0, 41, m * n)
    x = O.fill(0, 42, n)
    y0 = O.fill(0, 43, (m - 1) * incy + 1)
    dA, dx, dy = to_dev(A), to_dev(x), to_dev(y0)
    Q.dev_gemv(Q.MODES[mode], 1, m, n, 1.5, dA.data_ptr(), m, dx.data_ptr(), 1, -0.5, dy.data_ptr(), incy)
    cuda.cuda.synchronize()
    want = from_dev(dy)
    (hA, oA), (hx, ox), (hy, oy) = (_host(a, pinned, cuda) for a in (A, x, y0))
    with _mode(Q, mode):
        for _ in range(2):
            hy[...] = y0
            Q.gemv(1, m, n, 1.5, hA, m, hx, 1, -0.5, hy, incy)
            assert np.array_equal(hy, want)


@pytest.mark.parametrize("pinned", [False, True])
def test_host_dot_staged(O, Q, cuda, pinned):
    n, incy = 300001, 3
    _say(f"host dot pinned={pinned} n={n} incy={incy}")
    x = O.fill(0, 51, n)
    y = O.fill(0, 52, (n - 1) * incy + 1)
    T = Q.get_num_threads()
    dx, dy, dr = to_dev(x), to_dev(y), dev_empty(1)
    Q.dev_dot(Q.MODE_REFERENCE, n, dx.data_ptr(), 1, dy.data_ptr(), incy, T, dr.data_ptr())
    cuda.cuda.synchronize()
    want = pair(from_dev(dr)[0])
    assert want == O.to_pair(O.dot(n, x, 1, y, incy, T))
    (hx, ox), (hy, oy) = (_host(a, pinned, cuda) for a in (x, y))
    with _mode(Q, "reference"):
        for _ in range(2):
            assert pair(Q.dot(n, hx, 1, hy, incy)) == want


@pytest.mark.parametrize("pinned", [False, True])
def test_host_axpy_staged(O, Q, cuda, pinned):
    n, incx, incy = 400001, 1, 2
    _say(f"host axpy pinned={pinned} n={n} incy={incy}")
    x = O.fill(0, 61, n)
    y0 = O.fill(0, 62, (n - 1) * incy + 1)
    alpha = O.fill(0, 63, 1)[0]
    want = y0.copy()
    O.axpy(n, alpha, x, incx, want, incy)
    (hx, ox), (hy, oy) = (_host(a, pinned, cuda) for a in (x, y0))
    a = np.ascontiguousarray(alpha)
    for _ in range(2):
        hy[...] = y0
        Q.lib().quadblas_axpy_q(n, ctypes.c_void_p(a.ctypes.data), ctypes.c_void_p(hx.ctypes.data), incx,
                                ctypes.c_void_p(hy.ctypes.data), incy)
        assert Q.last_error() in ("", None)
        assert np.array_equal(hy, want)


def test_host_calls_from_two_threads(O, Q, cuda):
    # each host thread has its own stream, staging slots and workspaces: staged gemv (ColMajor,
    # 22 MiB of A, y staged both ways) and dot calls from two threads at once give the same bits
    # as the device-pointer routines
    import threading
    m, n, incy = 200003, 7, 2
    A = O.fill(0, 71, m * n)
    x = O.fill(0, 72, n)
    y0 = O.fill(0, 73, (m - 1) * incy + 1)
    dA, dx, dy = to_dev(A), to_dev(x), to_dev(y0)
    Q.dev_gemv(Q.MODE_DD, 1, m, n, 1.5, dA.data_ptr(), m, dx.data_ptr(), 1, -0.5, dy.data_ptr(), incy)
    cuda.cuda.synchronize()
    want_y = from_dev(dy)
    nd = 300001
    u, v = O.fill(0, 74, nd), O.fill(0, 75, nd)
    du, dv, dr = to_dev(u), to_dev(v), dev_empty(1)
    Q.dev_dot(Q.MODE_DD, nd, du.data_ptr(), 1, dv.data_ptr(), 1, 1, dr.data_ptr())
    cuda.cuda.synchronize()
    want_d = pair(from_dev(dr)[0])
    _say("host gemv and dot from two threads")
    errors = []

    def gemv_worker():
        try:
            for _ in range(4):
                hy = y0.copy()
                Q.gemv(1, m, n, 1.5, A, m, x, 1, -0.5, hy, incy)
                if not np.array_equal(hy, want_y):
                    errors.append("gemv")
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    def dot_worker():
        try:
            for _ in range(8):
                if pair(Q.dot(nd, u, 1, v, 1)) != want_d:
                    errors.append("dot")
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    with _mode(Q, "dd"):
        ts = [threading.Thread(target=gemv_worker), threading.Thread(target=dot_worker)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=100)
        assert not any(t.is_alive() for t in ts)
    assert errors == []
