"""GPU parity of the full-range path of the EXACT form and of the OZAKI form's exact repair.

The fast path keeps two 384-bit accumulation windows per lane, which hold products within about
2^124 of an anchor.  A dot, gemv row or GEMM entry with a product outside them is recomputed with
a fixed-point accumulator in LDS that spans every binary128 product (qblas_amd/csrc/wide.hpp).
The inputs here span up to the whole binary128 range: log-uniform over 2^+-6000, rows whose
products spread over 2^+-500, and mixes of subnormals and values near overflow.

Checks:
  - bit-exact against the oracle's arbitrary-width exact accumulator (EXACT dot, gemv in both
    layouts, batched gemv), every entry within 2^-104 of the exact value for the OZAKI form
    (slice and CRT methods, single and batched), status 0, and the count of items that took the
    full-range path (so the path under test is known to have run);
  - closed forms: 1 + 2^-113 is a tie that rounds to even (1) alone, up with an extra product
    of 2^-32000 and down with -2^-32000; a pair of products +-2^16000 that cancels exactly
    changes nothing.
"""
import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, to_dev

pytestmark = pytest.mark.gpu

BIAS = 16383
ONE = (0, 0x3FFF << 48)
ONE_ULP = (1, 0x3FFF << 48)        # 1 + 2^-112


def _say(msg):
    print(msg, flush=True)


def q_pow2(e, neg=False):
    """2^e exactly as binary128 bits (normal or subnormal), as a (2,) uint64 array."""
    s = (1 << 63) if neg else 0
    if e >= 1 - BIAS:
        return np.array([0, s | ((e + BIAS) << 48)], dtype=np.uint64)
    m = 1 << (e - (1 - BIAS - 112))            # subnormal: m * 2^(1 - BIAS - 112)
    assert 0 < m < (1 << 112)
    return np.array([m & ((1 << 64) - 1), s | (m >> 64)], dtype=np.uint64)


def rand_q(rng, n, e_lo, e_hi):
    """n values with random sign, leading-bit exponent uniform in [e_lo, e_hi] (subnormal below
    1 - BIAS, down to -16494) and random bits below the leading one."""
    out = np.zeros((n, 2), dtype=np.uint64)
    es = rng.integers(e_lo, e_hi + 1, n)
    for i in range(n):
        e = int(es[i])
        s = int(rng.integers(0, 2)) << 63
        f = int(rng.integers(0, 1 << 56)) << 56 | int(rng.integers(0, 1 << 56))   # 112 random bits
        if e >= 1 - BIAS:
            hi, lo = s | ((e + BIAS) << 48) | (f >> 64), f & ((1 << 64) - 1)
        else:
            d = (1 - BIAS) - e                                # 1..112
            m = (1 << (112 - d)) | (f >> d)
            hi, lo = s | (m >> 64), m & ((1 << 64) - 1)
        out[i] = (lo, hi)
    return out


def rand_q_pairs(rng, n):
    """x_i over the whole finite range (subnormals to near overflow), y_i such that each product's
    exponent stays at or below 2^15990 (the sum cannot overflow)."""
    ex = rng.integers(-16494, 16384, n)
    x = np.zeros((n, 2), dtype=np.uint64)
    y = np.zeros((n, 2), dtype=np.uint64)
    for i in range(n):
        e = int(ex[i])
        x[i] = rand_q(rng, 1, e, e)[0]
        hi_e = min(16383, 15990 - e)
        lo_e = -16494
        f = int(rng.integers(lo_e, hi_e + 1))
        y[i] = rand_q(rng, 1, f, f)[0]
    return x, y


def _dev_dot_exact(Q, cuda, n, x, incx=1, y=None, incy=1):
    dx, dy, dr = to_dev(x), to_dev(y), dev_empty(1)
    Q.dev_dot(Q.MODE_EXACT, n, dx.data_ptr(), incx, dy.data_ptr(), incy, 1, dr.data_ptr())
    cuda.cuda.synchronize()
    return pair(from_dev(dr)[0]), Q.dev_exact_wide_items(), Q.dev_exact_status()


# ------------------------------------------------------------------------------------------------
# dot
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [2, 5, 1000, 65537, 1 << 20])
def test_dot_wide_loguniform(O, Q, cuda, n):
    _say(f"exact dot, log-uniform 2^+-6000, n={n}")
    x = O.fill(1, 500 + n, n, -6000, 6000)
    y = O.fill(1, 501 + n, n, -6000, 6000)
    got, wide, st = _dev_dot_exact(Q, cuda, n, x, 1, y, 1)
    want, _ = O.exact_dot(x, y)
    assert st == 0
    assert got == want
    if n >= 1000:
        assert wide == 1


def test_dot_wide_strided_and_mixed_range(O, Q, cuda):
    rng = np.random.default_rng(7)
    n = 3000
    x, y = rand_q_pairs(rng, n)
    _say("exact dot, subnormals to near overflow, n=3000")
    got, wide, st = _dev_dot_exact(Q, cuda, n, x, 1, y, 1)
    want, _ = O.exact_dot(x, y)
    assert st == 0 and wide == 1 and got == want
    # the same values strided (incx = 3, incy = 2)
    xs = np.zeros((3 * n, 2), dtype=np.uint64)
    ys = np.zeros((2 * n, 2), dtype=np.uint64)
    xs[::3], ys[::2] = x, y
    _say("exact dot, mixed range, strided")
    got2, _, st2 = _dev_dot_exact(Q, cuda, n, xs, 3, ys, 2)
    assert st2 == 0 and got2 == want


def _tie_vectors(tiny_sign, huge):
    """x, y with sum x_i y_i = 1 + 2^-113 + tiny_sign * 2^-32000 (+ 2^16000 - 2^16000 if huge)."""
    xs, ys = [], []
    if huge:
        xs += [q_pow2(16000), q_pow2(16000)]
        ys += [q_pow2(0), q_pow2(0, neg=True)]
    xs += [q_pow2(0), q_pow2(-113)]
    ys += [q_pow2(0), q_pow2(0)]
    if tiny_sign:
        xs.append(q_pow2(-16400))                           # subnormal
        ys.append(q_pow2(-15600, neg=tiny_sign < 0))
    return np.array(xs), np.array(ys)


@pytest.mark.parametrize("huge", [False, True])
def test_dot_wide_tie_closed_form(O, Q, cuda, huge):
    for tiny, want in ((0, ONE), (1, ONE_ULP), (-1, ONE)):
        x, y = _tie_vectors(tiny, huge)
        _say(f"exact dot tie: tiny={tiny} huge={huge}")
        got, wide, st = _dev_dot_exact(Q, cuda, x.shape[0], x, 1, y, 1)
        assert st == 0
        assert got == want, (tiny, huge, got)
        assert got == O.exact_dot(x, y)[0]
        if tiny or huge:
            assert wide == 1
    # the same on a long vector: the tiny product and the huge pair far apart in the array
    n = 200003
    x = np.zeros((n, 2), dtype=np.uint64)
    y = np.zeros((n, 2), dtype=np.uint64)
    xv, yv = _tie_vectors(1, True)
    for slot, i in enumerate((5, 100000, 150001, 160000, n - 1)):
        x[i], y[i] = xv[slot], yv[slot]
    got, wide, st = _dev_dot_exact(Q, cuda, n, x, 1, y, 1)
    assert st == 0 and wide == 1 and got == ONE_ULP


# ------------------------------------------------------------------------------------------------
# gemv
# ------------------------------------------------------------------------------------------------

def _gemv_case(O, Q, cuda, layout, m, n, A, x, y, alpha, beta):
    lda = n if layout == 0 else m
    want = y.copy()
    O.exact_gemv(layout, m, n, alpha, A, lda, x, 1, beta, want, 1)
    dA, dx, dy = to_dev(A), to_dev(x), to_dev(y)
    Q.dev_gemv(Q.MODE_EXACT, layout, m, n, alpha, dA.data_ptr(), lda, dx.data_ptr(), 1, beta, dy.data_ptr(), 1)
    cuda.cuda.synchronize()
    got = from_dev(dy)
    return got, want, Q.dev_exact_wide_items(), Q.dev_exact_status()


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("spread", ["log6000", "rows500"])
def test_gemv_wide(O, Q, cuda, layout, spread):
    m, n = 200, 300
    if spread == "log6000":
        A = O.fill(1, 610 + layout, m * n, -6000, 6000)
        x = O.fill(1, 612, n, -6000, 6000)
    else:
        A = O.fill(1, 620 + layout, m * n, -250, 250)
        x = O.fill(1, 622, n, -250, 250)
    y = O.fill(0, 613, m)
    for alpha, beta in ((1.5, -0.5), (1.0, 0.0), ("q", "q"), (0.0, 2.0)):
        if alpha == "q":
            alpha, beta = O.fill(1, 614, 1, -16000, 16000)[0], O.fill(1, 615, 1, -100, 100)[0]
        _say(f"exact gemv wide layout={layout} spread={spread} alpha={alpha}")
        got, want, wide, st = _gemv_case(O, Q, cuda, layout, m, n, A, x, y, alpha, beta)
        assert st == 0
        assert np.array_equal(got, want), np.nonzero(np.any(got != want, axis=1))[0][:10]
        if not (isinstance(alpha, float) and alpha == 0.0):
            assert wide >= 1


def test_gemv_wide_tie_rows(O, Q, cuda):
    # x = [1, 1, 2^-15600, 2^16000, 2^16000]; row r = [1, 2^-113, t 2^-16400, h, -h] (t in
    # {0, 1, -1}, h in {0, 1}): the row's products are the tie vector's (1, 2^-113, t 2^-32000,
    # +-h 2^16000) and y_r (alpha = 1, beta = 0) is its closed form
    cases = [(0, 0, ONE), (1, 0, ONE_ULP), (-1, 0, ONE), (0, 1, ONE), (1, 1, ONE_ULP), (-1, 1, ONE)]
    m, n = len(cases), 5
    x = np.array([q_pow2(0), q_pow2(0), q_pow2(-15600), q_pow2(16000), q_pow2(16000)])
    A = np.zeros((m, n, 2), dtype=np.uint64)
    for r, (tiny, huge, _) in enumerate(cases):
        A[r, 0], A[r, 1] = q_pow2(0), q_pow2(-113)
        if tiny:
            A[r, 2] = q_pow2(-16400, neg=tiny < 0)
        if huge:
            A[r, 3], A[r, 4] = q_pow2(0), q_pow2(0, neg=True)
    y = np.zeros((m, 2), dtype=np.uint64)
    for layout in (0, 1):
        Am = A.reshape(m * n, 2) if layout == 0 else np.ascontiguousarray(A.transpose(1, 0, 2)).reshape(m * n, 2)
        _say(f"exact gemv tie rows layout={layout}")
        got, want_o, wide, st = _gemv_case(O, Q, cuda, layout, m, n, Am, x, y, 1.0, 0.0)
        assert st == 0
        assert [pair(g) for g in got] == [w for _, _, w in cases]
        assert np.array_equal(got, want_o)
        assert wide >= 4


# ------------------------------------------------------------------------------------------------
# OZAKI gemm: the exact repair on the full-range path
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("method", ["slices", "crt"])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("spread", [3000, 250])
def test_ozaki_wide(O, Q, cuda, method, layout, spread):
    m, n, k = 64, 48, 300
    A = O.fill(1, 700 + spread, m * k, -spread, spread)
    B = O.fill(1, 701 + spread, k * n, -spread, spread)
    C = O.fill(0, 702, m * n)
    if layout == 0:
        lda, ldb, ldc, As, Bs, Cs = k, n, n, (k, 1), (n, 1), (n, 1)
    else:
        lda, ldb, ldc, As, Bs, Cs = m, k, m, (1, m), (1, k), (1, m)
    prev = Q.get_ozaki_method()
    Q.set_ozaki_method(Q.OZAKI_SLICES if method == "slices" else Q.OZAKI_CRT)
    try:
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        _say(f"ozaki wide method={method} layout={layout} spread=2^+-{spread}")
        Q.dev_gemm(Q.MODE_OZAKI, layout, m, n, k, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.5, dC.data_ptr(), ldc)
        cuda.cuda.synchronize()
        st, wide, rep = Q.dev_exact_status(), Q.dev_exact_wide_items(), Q.dev_ozaki_repairs()
    finally:
        Q.set_ozaki_method(prev)
    got = from_dev(dC)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    cand = got[ii * Cs[0] + jj * Cs[1]]
    _, rel = O.exact_gemm_samples(ii, jj, k, 1.0, A, As, B, Bs, 0.5, C, Cs, cand=cand)
    assert st == 0
    assert rel.max() <= 2.0 ** -104, rel.max()
    assert rep >= 1 and wide >= 1, (rep, wide)


# ------------------------------------------------------------------------------------------------
# batched forms
# ------------------------------------------------------------------------------------------------

def test_gemv_batched_exact_wide(O, Q, cuda):
    from test_gpu_batched import Pool
    rng = np.random.default_rng(5)
    for layout in (0, 1):
        count = 12
        dims, arrays = [], []
        for p in range(count):
            m, n = int(rng.integers(20, 120)), int(rng.integers(20, 200))
            spread = (6000, 250, 20)[p % 3]      # every third problem stays on the fast path
            lda = n if layout == 0 else m
            arrays += [O.fill(1, 900 + 3 * p, m * n, -spread, spread), O.fill(1, 901 + 3 * p, n, -spread, spread),
                       O.fill(0, 902 + 3 * p, m)]
            ab = O.fill(0, 990 + p, 2)
            dims.append((m, n, lda, ab[0], ab[1]))
        pool = Pool(cuda, arrays)
        desc = np.zeros(count, Q.GEMV_PROBLEM)
        for p, (m, n, lda, a, b) in enumerate(dims):
            desc[p] = (m, n, lda, 1, 1, pool.ptr(3 * p), pool.ptr(3 * p + 1), pool.ptr(3 * p + 2), a, b)
        ddesc = cuda.from_numpy(desc.view(np.uint8)).to("cuda:0")
        _say(f"batched exact gemv wide layout={layout}")
        Q.dev_gemv_batched(Q.MODE_EXACT, layout, count, ddesc.data_ptr())
        cuda.cuda.synchronize()
        assert Q.dev_exact_status() == 0
        assert Q.dev_exact_wide_items() >= 1
        for p, (m, n, lda, a, b) in enumerate(dims):
            y = arrays[3 * p + 2].copy()
            O.exact_gemv(layout, m, n, a, arrays[3 * p], lda, arrays[3 * p + 1], 1, b, y, 1)
            assert np.array_equal(pool.get(3 * p + 2, m), y), (layout, p)


def test_gemm_batched_ozaki_wide(O, Q, cuda):
    from test_gpu_batched import Pool
    rng = np.random.default_rng(6)
    layout, count = 0, 6
    dims, arrays = [], []
    for p in range(count):
        m, n, k = (int(v) for v in rng.integers(8, 70, 3))
        spread = (3000, 250)[p % 2]
        arrays += [O.fill(1, 800 + 3 * p, m * k, -spread, spread), O.fill(1, 801 + 3 * p, k * n, -spread, spread),
                   O.fill(0, 802 + 3 * p, m * n)]
        ab = O.fill(0, 880 + p, 2)
        dims.append((m, n, k, ab[0], ab[1]))
    pool = Pool(cuda, arrays)
    desc = np.zeros(count, Q.GEMM_PROBLEM)
    for p, (m, n, k, a, b) in enumerate(dims):
        desc[p] = (m, n, k, k, n, n, pool.ptr(3 * p), pool.ptr(3 * p + 1), pool.ptr(3 * p + 2), a, b, 0)
    ddesc = cuda.from_numpy(desc.view(np.uint8)).to("cuda:0")
    _say("batched ozaki gemm wide")
    Q.dev_gemm_batched(Q.MODE_OZAKI, layout, count, ddesc.data_ptr())
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() == 0
    assert Q.dev_exact_wide_items() >= 1
    for p, (m, n, k, a, b) in enumerate(dims):
        got = pool.get(3 * p + 2, m * n)
        ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        ii, jj = ii.ravel(), jj.ravel()
        _, rel = O.exact_gemm_samples(ii, jj, k, a, arrays[3 * p], (k, 1), arrays[3 * p + 1], (n, 1), b,
                                      arrays[3 * p + 2], (n, 1), cand=got[ii * n + jj])
        assert rel.max() <= 2.0 ** -104, (p, rel.max())


# ------------------------------------------------------------------------------------------------
# through the C++ drop-in's entry points (Part 2): the global EXACT mode
# ------------------------------------------------------------------------------------------------

def test_dot_q_wide_through_host_abi(O, Q, cuda):
    import ctypes
    x = O.fill(1, 950, 4000, -500, 500)
    y = O.fill(1, 951, 4000, -500, 500)
    want, _ = O.exact_dot(x, y)
    prev = Q.get_mode()
    Q.set_mode(Q.MODE_EXACT)
    try:
        res = np.zeros(2, dtype=np.uint64)
        Q.lib().quadblas_dot_q(4000, x.ctypes.data_as(ctypes.c_void_p), 1, y.ctypes.data_as(ctypes.c_void_p), 1,
                               res.ctypes.data_as(ctypes.c_void_p))
        err = Q.last_error()
    finally:
        Q.set_mode(prev)
    assert err == ""
    assert pair(res) == want
