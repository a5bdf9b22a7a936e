"""GPU parity of the reference-order kernels: bit-for-bit against the oracle's restatement of the
reference's evaluation order, through the C ABI (device-pointer API and the reference's own
host-pointer entry points).

Edge cases follow the reference's tests and code paths: n = 0, tiny and odd n, the 500-element
parallel threshold (level1.hpp:40, level2.hpp:27), strided vectors, the 64 small-GEMM threshold
(level3.hpp:224-229), kc = 126 blocking, both layouts, alpha/beta variants, beta = 0 with NaN in y
(propagates, level2.hpp:48), leading dimensions larger than the matrix, and the transpose flags
of the C ABI.
"""
import math

import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, rows, to_dev

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_fill_matches_oracle(O, Q, cuda):
    for kind, seed, n in ((0, 1, 10007), (1, 2, 4096), (2, 3, 999), (3, 3, 999), (4, 3, 999)):
        _say(f"fill kind={kind} seed={seed} n={n}")
        d = dev_empty(n)
        Q.dev_fill(kind, seed, n, d.data_ptr(), -30, 30)
        cuda.cuda.synchronize()
        want = O.fill_pair_y(seed, n) if kind == 3 else O.fill(kind, seed, n, -30, 30)
        assert np.array_equal(from_dev(d), want)


DOT_CASES = [
    (0, 1, 1, 4), (1, 1, 1, 4), (2, 1, 1, 4), (7, 1, 1, 4), (499, 1, 1, 8), (500, 1, 1, 8), (501, 1, 1, 3),
    (1001, 1, 1, 1), (1001, 1, 1, 2), (4097, 1, 1, 13), (100001, 1, 1, 8), (100001, 1, 1, 64),
    (5, 2, 3, 4), (499, 2, 1, 8), (1000, 3, 2, 8), (4001, 1, 2, 7), (777, 0, 1, 4),
]


@pytest.mark.parametrize("n,incx,incy,T", DOT_CASES)
def test_dot_reference_order(O, Q, cuda, n, incx, incy, T):
    _say(f"dot n={n} incx={incx} incy={incy} T={T}")
    lx = max(1, (n - 1) * incx + 1)
    ly = max(1, (n - 1) * incy + 1)
    x = O.fill(0, 100 + n, lx)
    y = O.fill(0, 200 + n, ly)
    dx, dy, dr = to_dev(x), to_dev(y), dev_empty(1)
    Q.dev_dot(Q.MODE_REFERENCE, n, dx.data_ptr(), incx, dy.data_ptr(), incy, T, dr.data_ptr())
    cuda.cuda.synchronize()
    want = O.dot(n, x, incx, y, incy, T)
    assert pair(from_dev(dr)[0]) == O.to_pair(want)


def test_dot_cancellation_and_c_abi(O, Q, cuda):
    n = 1 << 16
    x = O.fill(2, 7, n)
    y = O.fill_pair_y(7, n)
    for T in (1, 8, 96):
        Q.set_num_threads(T)
        assert Q.get_num_threads() == T
        want = O.dot(n, x, 1, y, 1, T)
        got = Q.dot(n, x, 1, y, 1)                 # QuadBLAS::dot path (binary128 result)
        assert pair(got) == O.to_pair(want)
        d = Q.qdot(n, x, 1, y, 1)                  # quadblas_qdot returns double
        assert d == O.to_double(want)
    Q.set_num_threads(0)


def test_dot_known_answers(O, Q, cuda):
    # test_quadblas.cpp:205-224 and :715-739
    x = O.qarr(range(1, 11))
    one = O.qarr([1.0] * 10)
    assert Q.qdot(10, x, 1, one, 1) == 55.0
    x = O.qarr([1e20, 1.0, -1e20] + [0.0] * 7)
    assert pair(Q.dot(10, x, 1, one, 1)) == O.to_pair(O.q(1.0))
    assert Q.qdot(0, x, 1, one, 1) == 0.0


GEMV_CASES = [
    # layout, m, n, lda_pad, incx, incy, alpha, beta
    (0, 1, 1, 0, 1, 1, 1.0, 0.0), (0, 7, 5, 3, 1, 1, 1.5, -0.5), (0, 64, 65, 0, 1, 1, 1.0, 0.0),
    (0, 300, 301, 2, 2, 3, 1.5, -0.5), (0, 65, 130, 0, 1, 2, "q", "q"), (0, 500, 33, 0, 1, 1, 1.0, 1.0),
    (1, 1, 1, 0, 1, 1, 1.0, 0.0), (1, 7, 5, 3, 1, 1, 1.5, -0.5), (1, 64, 65, 0, 2, 1, 1.0, 0.0),
    (1, 300, 301, 1, 1, 3, "q", "q"), (1, 513, 17, 0, 1, 1, 1.5, -0.5),
]


@pytest.mark.parametrize("layout,m,n,pad,incx,incy,alpha,beta", GEMV_CASES)
def test_gemv_reference_order(O, Q, cuda, layout, m, n, pad, incx, incy, alpha, beta):
    _say(f"gemv layout={layout} m={m} n={n} pad={pad} incx={incx} incy={incy}")
    if alpha == "q":
        alpha, beta = O.fill(0, 31, 1)[0], O.fill(0, 32, 1)[0]
    lda = (n if layout == 0 else m) + pad
    nA = (m - 1) * lda + n if layout == 0 else (n - 1) * lda + m
    A = O.fill(0, 41 + m, nA)
    x = O.fill(0, 42 + n, (n - 1) * incx + 1)
    y = O.fill(0, 43, (m - 1) * incy + 1)
    want = y.copy()
    O.gemv(layout, m, n, alpha, A, lda, x, incx, beta, want, incy)
    dA, dx, dy = to_dev(A), to_dev(x), to_dev(y)
    Q.dev_gemv(Q.MODE_REFERENCE, layout, m, n, alpha, dA.data_ptr(), lda, dx.data_ptr(), incx, beta,
               dy.data_ptr(), incy)
    cuda.cuda.synchronize()
    assert np.array_equal(from_dev(dy), want)
    # the full-precision host entry point (QuadBLAS::gemv) gives the same bits
    got = y.copy()
    Q.gemv(layout, m, n, alpha, A, lda, x, incx, beta, got, incy)
    assert np.array_equal(got, want)


def test_gemv_c_abi_transpose_and_nan(O, Q, cuda):
    m, n = 37, 23
    A = O.fill(0, 51, m * n)       # row-major m x n
    x = O.fill(0, 52, m)
    y = O.fill(0, 53, n)
    # trans = 'T' on row-major (m x n): y (n) = A^T x  ==  ColMajor gemv with m, n swapped
    want = y.copy()
    O.gemv(1, n, m, 1.5, A, n, x, 1, -0.5, want, 1)
    got = y.copy()
    Q.qgemv("R", "T", m, n, 1.5, A, n, x, 1, -0.5, got, 1)
    assert np.array_equal(got, want)
    # layout 'C', trans 'N'
    want = O.fill(0, 54, m)
    got = want.copy()
    O.gemv(1, m, n, 2.0, A, m, y, 1, 0.25, want, 1)
    Q.qgemv("c", "n", m, n, 2.0, A, m, y, 1, 0.25, got, 1)
    assert np.array_equal(got, want)
    # beta = 0 still reads y: NaN propagates (level2.hpp:48)
    yn = O.fill(0, 55, m)
    yn[3] = pair(np.array([0, 0x7FFF800000000000], np.uint64))
    Q.qgemv("R", "N", m, n, 1.0, A, n, y, 1, 0.0, yn, 1)
    assert ((int(yn[3][1]) >> 48) & 0x7FFF) == 0x7FFF
    assert ((int(yn[2][1]) >> 48) & 0x7FFF) != 0x7FFF
    # known answer (test_quadblas.cpp:327-350)
    A = O.qarr(range(1, 10))
    x = O.qarr([1, 1, 1])
    y = O.qarr([0, 0, 0])
    Q.qgemv("R", "N", 3, 3, 1.0, A, 3, x, 1, 0.0, y, 1)
    assert list(O.to_double(y)) == [6.0, 15.0, 24.0]


GEMM_CASES = [
    # layout, m, n, k, pad, alpha, beta
    (0, 1, 1, 1, 0, 1.0, 0.0), (0, 2, 2, 2, 0, 1.0, 0.0), (0, 47, 31, 23, 0, 1.0, 0.0),
    (0, 64, 64, 64, 1, 2.5, 1.5), (0, 65, 3, 130, 0, 1.0, 0.0), (0, 20, 70, 253, 2, "q", "q"),
    (0, 100, 107, 103, 0, 1.5, -0.5), (0, 33, 17, 400, 0, 1.0, 1.0),
    (1, 1, 1, 1, 0, 1.0, 0.0), (1, 47, 31, 23, 0, 1.0, 0.0), (1, 65, 3, 130, 1, 1.5, -0.5),
    (1, 100, 107, 103, 0, "q", "q"), (1, 20, 70, 253, 0, 1.0, 0.0),
]


@pytest.mark.parametrize("layout,m,n,k,pad,alpha,beta", GEMM_CASES)
def test_gemm_reference_order(O, Q, cuda, layout, m, n, k, pad, alpha, beta):
    _say(f"gemm layout={layout} m={m} n={n} k={k} pad={pad}")
    if alpha == "q":
        alpha, beta = O.fill(0, 61, 1)[0], O.fill(0, 62, 1)[0]
    if layout == 0:
        lda, ldb, ldc = k + pad, n + pad, n + pad
        nA, nB, nC = (m - 1) * lda + k, (k - 1) * ldb + n, (m - 1) * ldc + n
    else:
        lda, ldb, ldc = m + pad, k + pad, m + pad
        nA, nB, nC = (k - 1) * lda + m, (n - 1) * ldb + k, (n - 1) * ldc + m
    A = O.fill(0, 71 + m, nA)
    B = O.fill(0, 72 + n, nB)
    C = O.fill(0, 73 + k, nC)
    want = C.copy()
    O.gemm(layout, m, n, k, alpha, A, lda, B, ldb, beta, want, ldc)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    Q.dev_gemm(Q.MODE_REFERENCE, layout, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta,
               dC.data_ptr(), ldc)
    cuda.cuda.synchronize()
    assert np.array_equal(from_dev(dC), want)


def test_gemm_c_abi(O, Q, cuda):
    # known answers (test_quadblas.cpp:446-464, :679-688) through quadblas_qgemm; trans ignored
    A = O.qarr([1, 2, 3, 4])
    B = O.qarr([5, 6, 7, 8])
    C = O.qarr([0, 0, 0, 0])
    Q.qgemm("R", "T", "T", 2, 2, 2, 1.0, A, 2, B, 2, 0.0, C, 2)
    assert list(O.to_double(C)) == [19.0, 22.0, 43.0, 50.0]
    m, n, k = 70, 66, 300
    A = O.fill(0, 81, m * k)
    B = O.fill(0, 82, k * n)
    C = O.fill(0, 83, m * n)
    want = C.copy()
    O.gemm(0, m, n, k, 1.5, A, k, B, n, -0.5, want, n)
    Q.qgemm("R", "N", "N", m, n, k, 1.5, A, k, B, n, -0.5, C, n)
    assert np.array_equal(C, want)


def test_degenerate_sizes_leave_outputs(O, Q, cuda):
    y = O.fill(0, 91, 5)
    y0 = y.copy()
    A = O.fill(0, 92, 25)
    Q.qgemv("R", "N", 0, 5, 1.0, A, 5, A, 1, 2.0, y, 1)
    Q.qgemv("R", "N", 5, 0, 1.0, A, 5, A, 1, 2.0, y, 1)
    Q.qgemm("R", "N", "N", 5, 5, 0, 1.0, A, 5, A, 5, 2.0, A.copy(), 5)
    assert np.array_equal(y, y0)


def test_config1_gemv_rowmajor_1500(O, Q, cuda):
    # The reference's own CPU benchmark case (README.md:23): plain qgemv RowMajor 1500 x 1500,
    # y := 1 A x + 0 y, entries a + b 2^-60 (full mantissa), y0 = 0, REFERENCE order: bit for bit
    # against the oracle's restatement of level2.hpp:15-50, through the device API and through the
    # reference's C ABI (quadblas_qgemv with host buffers).
    m = 1500
    A = O.fill(0, 42, m * m)
    x = O.fill(0, 43, m)
    want = np.zeros((m, 2), dtype=np.uint64)
    O.gemv(0, m, m, 1.0, A, m, x, 1, 0.0, want, 1)
    _say("config 1: qgemv RowMajor 1500^2, device API")
    dA, dx, dy = to_dev(A), to_dev(x), dev_empty(m)
    dy.zero_()
    Q.dev_gemv(Q.MODE_REFERENCE, 0, m, m, 1.0, dA.data_ptr(), m, dx.data_ptr(), 1, 0.0, dy.data_ptr(), 1)
    cuda.cuda.synchronize()
    assert np.array_equal(from_dev(dy), want)
    _say("config 1: qgemv RowMajor 1500^2, quadblas_qgemv (host buffers)")
    y = np.zeros((m, 2), dtype=np.uint64)
    prev = Q.get_mode()
    Q.set_mode(Q.MODE_REFERENCE)
    try:
        Q.qgemv("R", "N", m, m, 1.0, A, m, x, 1, 0.0, y, 1)
    finally:
        Q.set_mode(prev)
    assert np.array_equal(y, want)
