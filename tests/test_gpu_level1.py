"""GPU parity of qaxpy and qnrm2 (level1.hpp:139-223, c_interface.hpp:34-59, cpp_classes.hpp:72-81).

axpy: y_i = fma(alpha, x_i, y_i) per element, bit-exact against the oracle over sizes, strides and
n = 0.  nrm2: the dot of the mode (REFERENCE order with T, or EXACT) followed by the correctly rounded
square root, bit-exact against the oracle's dot and an exact integer square root.
"""
import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, sqrt_rn, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,incx,incy", [(0, 1, 1), (1, 1, 1), (7, 1, 1), (1000, 1, 1), (100003, 1, 1), (777, 2, 3),
                                         (5000, 3, 1)])
def test_axpy_bit_exact(O, Q, cuda, n, incx, incy):
    print(f"axpy n={n} incx={incx} incy={incy}", flush=True)
    x = O.fill(0, 3 + n, max(1, (n - 1) * incx + 1))
    y = O.fill(0, 4 + n, max(1, (n - 1) * incy + 1))
    alpha = O.fill(0, 5, 1)[0]
    want = y.copy()
    O.axpy(n, alpha, x, incx, want, incy)
    dx, dy = to_dev(x), to_dev(y)
    Q.dev_axpy(n, alpha, dx.data_ptr(), incx, dy.data_ptr(), incy)
    cuda.cuda.synchronize()
    assert np.array_equal(from_dev(dy), want)


@pytest.mark.parametrize("n,incx,T", [(1, 1, 4), (499, 1, 8), (10000, 1, 16), (100003, 2, 7)])
def test_nrm2_reference_and_exact(O, Q, cuda, n, incx, T):
    print(f"nrm2 n={n} incx={incx} T={T}", flush=True)
    x = O.fill(0, 9 + n, max(1, (n - 1) * incx + 1))
    dx, dr = to_dev(x), dev_empty(1)
    Q.dev_nrm2(Q.MODE_REFERENCE, n, dx.data_ptr(), incx, T, dr.data_ptr())
    cuda.cuda.synchronize()
    assert pair(from_dev(dr)[0]) == sqrt_rn(*O.to_pair(O.dot(n, x, incx, x, incx, T)))
    Q.dev_nrm2(Q.MODE_EXACT, n, dx.data_ptr(), incx, T, dr.data_ptr())
    cuda.cuda.synchronize()
    exact, _ = O.exact_dot(x, x, incx=incx, incy=incx, n=n)
    assert pair(from_dev(dr)[0]) == sqrt_rn(*pair(exact))


def test_level1_c_abi(O, Q, cuda):
    lib = Q.lib()
    x = O.qarr([3.0, 4.0])
    assert lib.quadblas_qnrm2(2, x.ctypes.data, 1) == 5.0
    assert lib.quadblas_qnrm2(0, x.ctypes.data, 1) == 0.0
    xs = O.qarr([1.0, 2.0, 3.0])
    ys = O.qarr([1.0, 1.0, 1.0])
    lib.quadblas_qaxpy(3, 2.0, xs.ctypes.data, 1, ys.ctypes.data, 1)
    assert list(O.to_double(ys)) == [3.0, 5.0, 7.0]
    lib.quadblas_qaxpy(0, 2.0, xs.ctypes.data, 1, ys.ctypes.data, 1)   # n <= 0: no-op
    assert list(O.to_double(ys)) == [3.0, 5.0, 7.0]
