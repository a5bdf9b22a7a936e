"""qblas_amd -- MI355X (gfx950) backend for the QuadBLAS binary128 hot path (dot, gemv, gemm).

The product is ``libqblas_hip.so`` (HIP kernels + C ABI, declared in ``include/qblas_hip.h``).  This
module is its host-side mirror for Python callers: thin ctypes bindings whose names, argument
order and meaning follow the reference's C interface (include/quadblas/interface/c_interface.hpp)
and its C++ routines (QuadBLAS::dot/gemv/gemm, include/quadblas/algorithms/level{1,2,3}.hpp).

Binary128 arrays are numpy ``uint64`` arrays of shape ``(..., 2)`` (``[lo, hi]`` words), i.e. the
memory image of ``__float128``/``Sleef_quad``.  Device-pointer functions take raw device addresses
(e.g. ``torch.Tensor.data_ptr()``) and a ``hipStream_t`` as an int (0 = default stream).

There is no CPU fallback: if the library is missing or a call fails, an exception is raised.
"""
import ctypes
import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqblas_hip.so")

MODE_REFERENCE = 0
MODE_EXACT = 1
MODE_DD = 2
MODE_OZAKI = 3
MODES = {"reference": MODE_REFERENCE, "exact": MODE_EXACT, "dd": MODE_DD, "ozaki": MODE_OZAKI}
OZAKI_AUTO = 0
OZAKI_SLICES = 1
OZAKI_CRT = 2

ROW_MAJOR = 0
COL_MAJOR = 1

_lib = None


def _bind_hip_runtime():
    """Make this process use one HIP runtime.

    PyTorch-ROCm wheels bundle their own libamdhip64 and reference it by the unversioned name
    ``libamdhip64.so``; libqblas_hip.so references ``libamdhip64.so.7`` (the runtime's soname).  If
    libqblas_hip.so were loaded first, a later ``import torch`` would map a second runtime into the
    process and the two would fight over the device.  Loading torch's copy first (without importing
    torch) makes both names resolve to it.  Without PyTorch the system runtime is used.
    """
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def lib():
    """Load libqblas_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m qblas_amd.build`")
    _bind_hip_runtime()
    L = ctypes.CDLL(LIB_PATH)
    P, I, I64, D, C = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_char
    sz = ctypes.c_size_t
    sig = {
        "quadblas_qdot": (D, [I, P, I, P, I]),
        "quadblas_qgemv": (None, [C, C, I, I, D, P, I, P, I, D, P, I]),
        "quadblas_qgemm": (None, [C, C, C, I, I, I, D, P, I, P, I, D, P, I]),
        "quadblas_qnrm2": (D, [I, P, I]),
        "quadblas_qaxpy": (None, [I, D, P, I, P, I]),
        "quadblas_axpy_q": (None, [sz, P, P, sz, P, sz]),
        "quadblas_nrm2_q": (None, [sz, P, sz, P]),
        "qblas_hip_axpy": (I, [I64, P, P, I64, P, I64, P]),
        "qblas_hip_nrm2": (I, [I, I64, P, I64, I, P, P]),
        "quadblas_set_num_threads": (None, [I]),
        "quadblas_get_num_threads": (I, []),
        "quadblas_get_version": (ctypes.c_char_p, []),
        "quadblas_is_aligned": (I, [P]),
        "quadblas_dot_q": (None, [sz, P, sz, P, sz, P]),
        "quadblas_gemv_q": (None, [I, sz, sz, P, P, sz, P, sz, P, P, sz]),
        "quadblas_gemm_q": (None, [I, sz, sz, sz, P, P, sz, P, sz, P, P, sz]),
        "qblas_set_mode": (I, [I]),
        "qblas_get_mode": (I, []),
        "qblas_set_kc": (None, [I]),
        "qblas_get_kc": (I, []),
        "quadblas_hip_last_error": (ctypes.c_char_p, []),
        "qblas_clear_error": (None, []),
        "qblas_hip_dot": (I, [I, I64, P, I64, P, I64, I, P, P]),
        "qblas_hip_gemv": (I, [I, I, I64, I64, P, P, I64, P, I64, P, P, I64, P]),
        "qblas_hip_gemm": (I, [I, I, I64, I64, I64, P, P, I64, P, I64, P, P, I64, P]),
        "qblas_hip_fill": (I, [I, ctypes.c_uint64, I64, I, I, P, P]),
        "qblas_hip_fill_2d": (I, [I, ctypes.c_uint64, I64, I64, I64, I, I, I, P, P]),
        "qblas_hip_exact_status": (I, [P]),
        "qblas_hip_exact_wide_items": (ctypes.c_longlong, [P]),
        "qblas_set_ozaki_slices": (I, [I]),
        "qblas_get_ozaki_slices": (I, []),
        "qblas_hip_ozaki_repairs": (ctypes.c_longlong, [P]),
        "qblas_hip_ozaki_last_slices": (I, []),
        "qblas_set_ozaki_method": (I, [I]),
        "qblas_get_ozaki_method": (I, []),
        "qblas_set_ozaki_crt_bits": (I, [I]),
        "qblas_get_ozaki_crt_bits": (I, []),
        "qblas_hip_ozaki_last_method": (I, []),
        "qblas_hip_ozaki_last_crt_bits": (I, []),
        "qblas_hip_ozaki_last_moduli": (I, []),
        "qblas_hip_set_phase_timing": (None, [I]),
        "qblas_hip_gemm_batched": (I, [I, I, I64, P, P]),
        "qblas_hip_gemv_batched": (I, [I, I, I64, P, P]),
        "qblas_hip_gemm_batched_tune": (I, [I, I, I64, P, P]),
        "qblas_hip_gemm_batched_config": (ctypes.c_char_p, [I, I]),
        "qblas_hip_gemm_batched_select": (I, [I, I, I64, I]),
        "qblas_hip_phase_times": (I, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().quadblas_hip_last_error().decode()


def _raise_if_error(what):
    err = last_error()
    if err:
        lib().qblas_clear_error()
        raise RuntimeError(f"{what}: {err}")


def _check(rc, what):
    if rc != 0:
        err = last_error()
        lib().qblas_clear_error()
        raise RuntimeError(f"{what} failed (status {rc}): {err}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _q(v):
    """binary128 scalar as a 16-byte buffer: accepts a (lo, hi) pair / (2,) array, or a float
    (converted exactly through the library's double->binary128 rule)."""
    buf = np.zeros(2, dtype=np.uint64)
    if isinstance(v, (tuple, list, np.ndarray)):
        buf[0], buf[1] = int(v[0]), int(v[1])
    else:
        buf[:] = from_double(float(v))
    return buf


def from_double(d):
    """Exact double -> binary128 bits (lo, hi) (IEEE: 52-bit fraction shifted into 112 bits)."""
    u = int(np.array([d], dtype=np.float64).view(np.uint64)[0])
    s, e, f = u >> 63, (u >> 52) & 0x7FF, u & ((1 << 52) - 1)
    if e == 0x7FF:
        hi = (s << 63) | (0x7FFF << 48) | ((1 << 47) if f else 0)
        return np.array([0, hi], dtype=np.uint64)
    if e == 0:
        if f == 0:
            return np.array([0, s << 63], dtype=np.uint64)
        lz = 52 - f.bit_length()
        f = (f << (lz + 1)) & ((1 << 52) - 1)
        e = 1 - lz - 1
        be = e - 1023 + 16383
    else:
        be = e - 1023 + 16383
    frac = f << 60
    return np.array([frac & ((1 << 64) - 1), (s << 63) | (be << 48) | (frac >> 64)], dtype=np.uint64)


def _arr(a):
    assert isinstance(a, np.ndarray) and a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], \
        "binary128 arrays are C-contiguous uint64 (..., 2)"
    return a


# ------------------------------------------------------------------------------------------------
# reference C ABI (host pointers)
# ------------------------------------------------------------------------------------------------

def qdot(n, x, incx, y, incy):
    """quadblas_qdot (c_interface.hpp:21-31): binary128 dot rounded to double."""
    r = lib().quadblas_qdot(n, _ptr(_arr(x)), incx, _ptr(_arr(y)), incy)
    _raise_if_error("quadblas_qdot")
    return r


def qgemv(layout, trans, m, n, alpha, A, lda, x, incx, beta, y, incy):
    """quadblas_qgemv (c_interface.hpp:64-90); y updated in place."""
    lib().quadblas_qgemv(layout.encode() if isinstance(layout, str) else layout,
                         trans.encode() if isinstance(trans, str) else trans, m, n, float(alpha),
                         _ptr(_arr(A)), lda, _ptr(_arr(x)), incx, float(beta), _ptr(_arr(y)), incy)
    _raise_if_error("quadblas_qgemv")


def qgemm(layout, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """quadblas_qgemm (c_interface.hpp:95-117); C updated in place; trans flags ignored."""
    enc = lambda c: c.encode() if isinstance(c, str) else c
    lib().quadblas_qgemm(enc(layout), enc(transa), enc(transb), m, n, k, float(alpha), _ptr(_arr(A)), lda,
                         _ptr(_arr(B)), ldb, float(beta), _ptr(_arr(C)), ldc)
    _raise_if_error("quadblas_qgemm")


def set_num_threads(t):
    lib().quadblas_set_num_threads(int(t))


def get_num_threads():
    return lib().quadblas_get_num_threads()


def get_version():
    return lib().quadblas_get_version().decode()


def is_aligned(addr):
    return lib().quadblas_is_aligned(ctypes.c_void_p(addr))


# ------------------------------------------------------------------------------------------------
# full-precision routines (QuadBLAS::dot / gemv / gemm)
# ------------------------------------------------------------------------------------------------

def dot(n, x, incx, y, incy):
    """QuadBLAS::dot (level1.hpp:80-137): binary128 result as a (2,) uint64 array."""
    out = np.zeros(2, dtype=np.uint64)
    lib().quadblas_dot_q(n, _ptr(_arr(x)), incx, _ptr(_arr(y)), incy, _ptr(out))
    _raise_if_error("quadblas_dot_q")
    return out


def gemv(layout, m, n, alpha, A, lda, x, incx, beta, y, incy):
    """QuadBLAS::gemv (level2.hpp:85-99); layout 0 = RowMajor, 1 = ColMajor; y in place."""
    a, b = _q(alpha), _q(beta)
    lib().quadblas_gemv_q(int(layout), m, n, _ptr(a), _ptr(_arr(A)), lda, _ptr(_arr(x)), incx, _ptr(b),
                          _ptr(_arr(y)), incy)
    _raise_if_error("quadblas_gemv_q")


def gemm(layout, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """QuadBLAS::gemm (level3.hpp:215-336); C in place."""
    a, b = _q(alpha), _q(beta)
    lib().quadblas_gemm_q(int(layout), m, n, k, _ptr(a), _ptr(_arr(A)), lda, _ptr(_arr(B)), ldb, _ptr(b),
                          _ptr(_arr(C)), ldc)
    _raise_if_error("quadblas_gemm_q")


def set_mode(mode):
    mode = MODES.get(mode, mode)
    _check(lib().qblas_set_mode(int(mode)), "qblas_set_mode")


def get_mode():
    return lib().qblas_get_mode()


def set_kc(kc):
    lib().qblas_set_kc(int(kc))


def get_kc():
    return lib().qblas_get_kc()


def set_ozaki_slices(s):
    """Slice count of the OZAKI gemm form (8..28), 0 = chosen by the library."""
    _check(lib().qblas_set_ozaki_slices(int(s)), "qblas_set_ozaki_slices")


def get_ozaki_slices():
    return lib().qblas_get_ozaki_slices()


def set_ozaki_method(method):
    """OZAKI_AUTO (default), OZAKI_SLICES or OZAKI_CRT."""
    _check(lib().qblas_set_ozaki_method(int(method)), "qblas_set_ozaki_method")


def get_ozaki_method():
    return lib().qblas_get_ozaki_method()


def set_ozaki_crt_bits(t):
    """Bits per scaled operand of the CRT method (48..158), 0 = chosen by the library."""
    _check(lib().qblas_set_ozaki_crt_bits(int(t)), "qblas_set_ozaki_crt_bits")


def get_ozaki_crt_bits():
    return lib().qblas_get_ozaki_crt_bits()


# ------------------------------------------------------------------------------------------------
# device-pointer API
# ------------------------------------------------------------------------------------------------

def dev_dot(mode, n, x, incx, y, incy, T, result, stream=0):
    _check(lib().qblas_hip_dot(int(mode), n, x, incx, y, incy, int(T), result, stream), "qblas_hip_dot")


def dev_gemv(mode, layout, m, n, alpha, A, lda, x, incx, beta, y, incy, stream=0):
    a, b = _q(alpha), _q(beta)
    _check(lib().qblas_hip_gemv(int(mode), int(layout), m, n, _ptr(a), A, lda, x, incx, _ptr(b), y, incy, stream),
           "qblas_hip_gemv")


def dev_gemm(mode, layout, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=0):
    a, b = _q(alpha), _q(beta)
    _check(lib().qblas_hip_gemm(int(mode), int(layout), m, n, k, _ptr(a), A, lda, B, ldb, _ptr(b), C, ldc, stream),
           "qblas_hip_gemm")


def dev_exact_status(stream=0):
    """1 if a term of the last EXACT/OZAKI-form device call of this thread fell outside the
    full-range accumulator (output NaN; cannot happen for binary128 inputs), 0 if the result is
    exact, -1 on an error (including a workspace reused by a later call on the stream, and a graph
    replay of a batched OZAKI call whose descriptors no longer have the captured shapes)."""
    return lib().qblas_hip_exact_status(stream)


def dev_exact_wide_items(stream=0):
    """Items (dots, rows, GEMM entries) the last EXACT- or OZAKI-form call on this thread computed
    on the full-range path (products too far apart for the accumulation windows); -1 if unknown."""
    return lib().qblas_hip_exact_wide_items(stream)


def dev_ozaki_repairs(stream=0):
    """Entries of the last OZAKI gemm of this thread that were recomputed exactly (-1: none)."""
    return lib().qblas_hip_ozaki_repairs(stream)


def ozaki_last_slices():
    """Slice count used by the last OZAKI gemm of this thread (0: none, or the CRT method)."""
    return lib().qblas_hip_ozaki_last_slices()


def ozaki_last_method():
    """OZAKI_SLICES or OZAKI_CRT for the last OZAKI gemm of this thread (0: none)."""
    return lib().qblas_hip_ozaki_last_method()


def ozaki_last_crt():
    """(bits, moduli) of the last OZAKI gemm of this thread if it used the CRT method, else (0, 0)."""
    return lib().qblas_hip_ozaki_last_crt_bits(), lib().qblas_hip_ozaki_last_moduli()


OZAKI_PHASES = ("slicing", "products", "finish", "repair")


def set_phase_timing(on):
    """Record HIP events between the phases of OZAKI gemm calls (on the call's stream)."""
    lib().qblas_hip_set_phase_timing(1 if on else 0)


def phase_times():
    """{phase: ms} of the last OZAKI gemm of this thread (synchronizes with it), or None."""
    buf = (ctypes.c_float * 4)()
    if lib().qblas_hip_phase_times(ctypes.cast(buf, ctypes.c_void_p)) != 4:
        return None
    return dict(zip(OZAKI_PHASES, [float(v) for v in buf]))


# ------------------------------------------------------------------------------------------------
# batched problems: descriptor arrays with the layout of qblas_gemm_problem / qblas_gemv_problem
# ------------------------------------------------------------------------------------------------
GEMM_PROBLEM = np.dtype([("m", "<i8"), ("n", "<i8"), ("k", "<i8"), ("lda", "<i8"), ("ldb", "<i8"), ("ldc", "<i8"),
                         ("A", "<u8"), ("B", "<u8"), ("C", "<u8"), ("alpha", "<u8", (2,)), ("beta", "<u8", (2,)),
                         ("reserved", "<i8")])
GEMV_PROBLEM = np.dtype([("m", "<i8"), ("n", "<i8"), ("lda", "<i8"), ("incx", "<i8"), ("incy", "<i8"),
                         ("A", "<u8"), ("x", "<u8"), ("y", "<u8"), ("alpha", "<u8", (2,)), ("beta", "<u8", (2,))])
assert GEMM_PROBLEM.itemsize == 112 and GEMV_PROBLEM.itemsize == 96


def dev_gemm_batched(mode, layout, count, problems, stream=0):
    """problems: device pointer to count qblas_gemm_problem descriptors."""
    _check(lib().qblas_hip_gemm_batched(int(mode), int(layout), count, problems, stream), "qblas_hip_gemm_batched")


def dev_gemv_batched(mode, layout, count, problems, stream=0):
    _check(lib().qblas_hip_gemv_batched(int(mode), int(layout), count, problems, stream), "qblas_hip_gemv_batched")


def dev_gemm_batched_tune(mode, layout, count, problems, stream=0):
    """Time every tile configuration on a replica of the batch; returns the chosen index."""
    r = lib().qblas_hip_gemm_batched_tune(int(mode), int(layout), count, problems, stream)
    if r < 0:
        raise RuntimeError(f"qblas_hip_gemm_batched_tune: status {-r}: {last_error()}")
    return r


def gemm_batched_config(mode, idx):
    s = lib().qblas_hip_gemm_batched_config(int(mode), int(idx))
    return s.decode() if s else None


def gemm_batched_select(mode, layout, count, idx):
    _check(lib().qblas_hip_gemm_batched_select(int(mode), int(layout), count, int(idx)), "qblas_hip_gemm_batched_select")


FILL_UNIFORM_FULL = 0
FILL_LOG_UNIFORM = 1
FILL_CANCEL_X = 2
FILL_CANCEL_Y = 3
FILL_CANCEL_X22 = 4   # FILL_CANCEL_X with +-(1e22*u) pairs


def dev_fill(kind, seed, n, out, lo_exp=-30, hi_exp=30, stream=0):
    _check(lib().qblas_hip_fill(int(kind), int(seed), n, int(lo_exp), int(hi_exp), out, stream), "qblas_hip_fill")


def dev_fill_2d(kind, seed, rows, cols, ld, layout, out, lo_exp=-30, hi_exp=30, stream=0):
    """Generator value r*cols + c stored at (r, c) of a rows x cols matrix in the given layout."""
    _check(lib().qblas_hip_fill_2d(int(kind), int(seed), rows, cols, ld, int(layout), int(lo_exp), int(hi_exp), out,
                                   stream), "qblas_hip_fill_2d")


def dev_axpy(n, alpha, x, incx, y, incy, stream=0):
    a = _q(alpha)
    _check(lib().qblas_hip_axpy(n, _ptr(a), x, incx, y, incy, stream), "qblas_hip_axpy")


def dev_nrm2(mode, n, x, incx, T, result, stream=0):
    _check(lib().qblas_hip_nrm2(int(mode), n, x, incx, int(T), result, stream), "qblas_hip_nrm2")
