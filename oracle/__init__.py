"""CPU oracle for the QuadBLAS binary128 hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this
package, and only as the checker (or, for the baseline, as the timed CPU restatement).  The
shipped library (``qblas_amd``) never imports it.

Binary128 arrays are numpy ``uint64`` arrays of shape ``(..., 2)`` holding ``[lo, hi]`` words,
i.e. exactly the little-endian memory image of ``__float128`` / ``Sleef_quad``.

Functions map one-to-one onto ``oracle/qoracle.c``; every restatement there cites the reference
file:line it follows (include/quadblas/algorithms/level{1,2,3}.hpp, detail/blocking.hpp).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


class Q128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


def build():
    """Compile oracle/qoracle.c into oracle/liboracle.so (gcc, no fast-math)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        sz = ctypes.c_size_t
        L.or_add.restype = Q128
        L.or_add.argtypes = [Q128, Q128]
        L.or_mul.restype = Q128
        L.or_mul.argtypes = [Q128, Q128]
        L.or_fma.restype = Q128
        L.or_fma.argtypes = [Q128, Q128, Q128]
        L.or_fmaq.restype = Q128
        L.or_fmaq.argtypes = [Q128, Q128, Q128]
        L.or_from_double.restype = Q128
        L.or_from_double.argtypes = [ctypes.c_double]
        L.or_to_double.restype = ctypes.c_double
        L.or_to_double.argtypes = [Q128]
        L.or_fma_selftest.restype = ctypes.c_uint64
        L.or_fma_selftest.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P]
        L.or_use_fmaq.argtypes = [ctypes.c_int]
        L.or_dot_kernel.restype = Q128
        L.or_dot_kernel.argtypes = [P, P, sz]
        L.or_dot.restype = Q128
        L.or_dot.argtypes = [sz, P, sz, P, sz, ctypes.c_int]
        L.or_gemv.argtypes = [ctypes.c_int, sz, sz, Q128, P, sz, P, sz, Q128, P, sz]
        L.or_gemm.argtypes = [ctypes.c_int, sz, sz, sz, Q128, P, sz, P, sz, Q128, P, sz, sz]
        L.or_blocking_kc.restype = sz
        L.or_blocking_kc.argtypes = [sz]
        L.or_exact_dot.restype = ctypes.c_int
        L.or_exact_dot.argtypes = [sz, P, ctypes.c_ssize_t, P, ctypes.c_ssize_t, Q128, Q128, Q128,
                                   P, sz, P, P]
        L.or_exact_gemm_samples.restype = ctypes.c_int
        L.or_exact_gemm_samples.argtypes = [sz, P, P, sz, Q128, P, ctypes.c_int64, ctypes.c_int64,
                                            P, ctypes.c_int64, ctypes.c_int64, Q128, P,
                                            ctypes.c_int64, ctypes.c_int64, P, P, P]
        L.or_exact_gemv.restype = ctypes.c_int
        L.or_exact_gemv.argtypes = [ctypes.c_int, sz, sz, Q128, P, sz, P, sz, Q128, P, sz]
        L.or_axpy.argtypes = [sz, Q128, P, sz, P, sz]
        L.or_rand64.restype = ctypes.c_uint64
        L.or_rand64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.or_fill.argtypes = [ctypes.c_int, ctypes.c_uint64, sz, ctypes.c_int, ctypes.c_int, P]
        L.or_fill_range.argtypes = [ctypes.c_int, ctypes.c_uint64, sz, sz, ctypes.c_int, ctypes.c_int, P]
        L.or_fill_pair_y.argtypes = [ctypes.c_uint64, sz, P]
        L.or_q2d.restype = ctypes.c_double
        L.or_q2d.argtypes = [Q128]
        _lib = L
    return _lib


# ------------------------------------------------------------------------------------------------
# value helpers
# ------------------------------------------------------------------------------------------------

def q(v):
    """A Q128 from a Python float/int (exactly representable in double), a (lo, hi) pair, or a
    length-2 uint64 array."""
    if isinstance(v, Q128):
        return v
    if isinstance(v, (tuple, list)) or (isinstance(v, np.ndarray) and v.shape == (2,)):
        return Q128(int(v[0]), int(v[1]))
    return lib().or_from_double(float(v))


def qarr(vals):
    """uint64 (n, 2) array from a sequence of doubles."""
    vals = list(vals)
    out = np.zeros((len(vals), 2), dtype=np.uint64)
    for i, v in enumerate(vals):
        r = lib().or_from_double(float(v))
        out[i, 0], out[i, 1] = r.lo, r.hi
    return out


def to_pair(r):
    return (int(r.lo), int(r.hi))


def to_double(a):
    """binary128 -> double (round to nearest even) for a Q128 or an (n, 2) array."""
    L = lib()
    if isinstance(a, Q128):
        return L.or_q2d(a)
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2)
    return np.array([L.or_q2d(Q128(int(r[0]), int(r[1]))) for r in a])


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _chk(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"] and a.shape[-1] == 2
    return a


# ------------------------------------------------------------------------------------------------
# arithmetic
# ------------------------------------------------------------------------------------------------

def add(a, b):
    return lib().or_add(q(a), q(b))


def mul(a, b):
    return lib().or_mul(q(a), q(b))


def fma(a, b, c):
    return lib().or_fma(q(a), q(b), q(c))


def fmaq(a, b, c):
    return lib().or_fmaq(q(a), q(b), q(c))


def fma_selftest(seed, count, mode):
    first = (Q128 * 5)()
    bad = lib().or_fma_selftest(seed, count, mode, ctypes.cast(first, ctypes.c_void_p))
    return int(bad), [to_pair(first[i]) for i in range(5)]


class use_fmaq:
    """Context manager: run the restatements on libquadmath fmaq instead of the oracle's fma."""

    def __enter__(self):
        lib().or_use_fmaq(1)

    def __exit__(self, *a):
        lib().or_use_fmaq(0)


# ------------------------------------------------------------------------------------------------
# restatements of the reference routines
# ------------------------------------------------------------------------------------------------

def dot_kernel(x, y):
    """level1.hpp:14-35 (2-lane even/odd chains + tail)."""
    _chk(x), _chk(y)
    return lib().or_dot_kernel(_ptr(x), _ptr(y), x.shape[0])


def dot(n, x, incx, y, incy, T):
    """level1.hpp:80-137 with T = omp_get_max_threads()."""
    _chk(x), _chk(y)
    return lib().or_dot(n, _ptr(x), incx, _ptr(y), incy, int(T))


def axpy(n, alpha, x, incx, y, incy):
    """level1.hpp:139-223; updates y in place."""
    _chk(x), _chk(y)
    lib().or_axpy(n, q(alpha), _ptr(x), incx, _ptr(y), incy)


def gemv(colmajor, m, n, alpha, A, lda, x, incx, beta, y, incy):
    """level2.hpp:15-99; updates y (uint64 (., 2) array) in place."""
    _chk(A), _chk(x), _chk(y)
    lib().or_gemv(int(colmajor), m, n, q(alpha), _ptr(A), lda, _ptr(x), incx, q(beta), _ptr(y), incy)


def gemm(colmajor, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kc=0):
    """level3.hpp:187-336 (kc = 126 on x86-64 unless overridden); updates C in place."""
    _chk(A), _chk(B), _chk(C)
    lib().or_gemm(int(colmajor), m, n, k, q(alpha), _ptr(A), lda, _ptr(B), ldb, q(beta), _ptr(C), ldc, kc)


def blocking_kc(k):
    return lib().or_blocking_kc(k)


# ------------------------------------------------------------------------------------------------
# exact accumulation
# ------------------------------------------------------------------------------------------------

def exact_dot(x, y, incx=1, incy=1, n=None, alpha=1.0, beta=0.0, c=0.0, cand=None):
    """Exact alpha*sum(x*y) + beta*c rounded once; returns (rounded Q128 pair, relerr array)."""
    _chk(x), _chk(y)
    if n is None:
        n = x.shape[0]
    cands = np.zeros((0, 2), np.uint64) if cand is None else np.ascontiguousarray(cand, np.uint64).reshape(-1, 2)
    rounded = Q128()
    rel = np.zeros(cands.shape[0], dtype=np.float64)
    rc = lib().or_exact_dot(n, _ptr(x), incx, _ptr(y), incy, q(alpha), q(beta), q(c),
                            _ptr(cands), cands.shape[0], ctypes.byref(rounded), _ptr(rel))
    if rc != 0:
        raise ValueError("exact_dot needs finite inputs")
    return to_pair(rounded), rel


def exact_gemm_samples(ii, jj, k, alpha, A, a_strides, B, b_strides, beta, C0, c_strides, cand=None):
    """Exact sampled GEMM entries.  a_strides = (row stride, col stride) in elements, etc."""
    ii = np.ascontiguousarray(ii, dtype=np.int64)
    jj = np.ascontiguousarray(jj, dtype=np.int64)
    ns = ii.shape[0]
    rounded = np.zeros((ns, 2), dtype=np.uint64)
    rel = np.zeros(ns, dtype=np.float64)
    c0p = _ptr(C0) if C0 is not None else None
    cp = _ptr(np.ascontiguousarray(cand, np.uint64)) if cand is not None else None
    if cand is not None:
        cand = np.ascontiguousarray(cand, np.uint64)
        cp = _ptr(cand)
    rc = lib().or_exact_gemm_samples(ns, _ptr(ii), _ptr(jj), k, q(alpha), _ptr(A), a_strides[0], a_strides[1],
                                     _ptr(B), b_strides[0], b_strides[1], q(beta), c0p,
                                     c_strides[0], c_strides[1], cp, _ptr(rounded),
                                     _ptr(rel) if cand is not None else None)
    if rc != 0:
        raise ValueError("exact_gemm_samples needs finite inputs")
    return rounded, rel


def exact_gemv(colmajor, m, n, alpha, A, lda, x, incx, beta, y, incy):
    """y <- RN(alpha*A x + beta*y) row by row, each row rounded once (in place)."""
    _chk(A), _chk(x), _chk(y)
    if lib().or_exact_gemv(int(colmajor), m, n, q(alpha), _ptr(A), lda, _ptr(x), incx, q(beta), _ptr(y), incy) != 0:
        raise ValueError("exact_gemv needs finite inputs")


def abs_array(a):
    """|a| elementwise (sign bits cleared)."""
    out = np.array(a, dtype=np.uint64, copy=True)
    out[..., 1] &= np.uint64((1 << 63) - 1)
    return out


# ------------------------------------------------------------------------------------------------
# generators
# ------------------------------------------------------------------------------------------------

KIND_UNIFORM_FULL = 0   # a + b*2^-60, a, b uniform [-1, 1) doubles, summed in binary128
KIND_LOG_UNIFORM = 1    # sign +-, exponent uniform in [lo, hi), 112 random fraction bits
KIND_CANCEL_PAIRS = 2   # +-(1e20*u) pairs on every 8-block's first two slots, kind 0 elsewhere
KIND_CANCEL_PAIRS_22 = 4  # the same with +-(1e22*u) pairs


def fill(kind, seed, n, lo_exp=-30, hi_exp=30, start=0):
    """Values start .. start + n - 1 of the seeded sequence of the given kind."""
    out = np.empty((n, 2), dtype=np.uint64)
    lib().or_fill_range(kind, seed, start, n, lo_exp, hi_exp, _ptr(out))
    return out


def fill_pair_y(seed, n):
    out = np.empty((n, 2), dtype=np.uint64)
    lib().or_fill_pair_y(seed, n, _ptr(out))
    return out


def rand64(seed, idx):
    return lib().or_rand64(seed, idx)
