"""GPU parity of the fast forms of dot and gemv.

EXACT form: bit-for-bit against the oracle's exact accumulator (the correctly rounded value of
alpha*sum(a*b) + beta*y), at small sizes across edge cases and at the full benchmark sizes
(qdot n = 2^24, qgemv 8192 x 8192), plus size-independent properties of a correctly rounded
result: invariance under swapping the operands, under permutation, and RowMajor == ColMajor on
transposed storage.

DD form (double-double on FP64): within the stated bound
    |result - exact| <= 2^-100 * (|alpha| * sum |a_j b_j| + |beta y|) + 1 ulp(result)
(inputs carry 2^-107 representation error; products/sums are error-free transforms).
"""
import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, to_dev, to_fraction

pytestmark = pytest.mark.gpu

DD_TOL = 2.0 ** -100


def _say(msg):
    print(msg, flush=True)


def _vecs(O, kind, n, seed, incx=1, incy=1):
    lx, ly = max(1, (n - 1) * incx + 1), max(1, (n - 1) * incy + 1)
    if kind == "cancel":
        x = O.fill(2, seed, lx)
        y = O.fill_pair_y(seed, ly)
    elif kind == "log":
        x = O.fill(1, seed, lx, -20, 20)
        y = O.fill(1, seed + 1, ly, -20, 20)
    else:
        x = O.fill(0, seed, lx)
        y = O.fill(0, seed + 1, ly)
    return x, y


def _dev_dot(Q, cuda, mode, n, x, incx, y, incy):
    dx, dy, dr = to_dev(x), to_dev(y), dev_empty(1)
    Q.dev_dot(mode, n, dx.data_ptr(), incx, dy.data_ptr(), incy, 1, dr.data_ptr())
    cuda.cuda.synchronize()
    return pair(from_dev(dr)[0])


DOT_CASES = [(1, 1, 1, "full"), (2, 1, 1, "full"), (63, 1, 1, "full"), (64, 1, 1, "cancel"),
             (1000, 1, 1, "full"), (4097, 1, 1, "log"), (100003, 1, 1, "cancel"), (100003, 1, 1, "full"),
             (777, 2, 3, "full"), (5000, 1, 0, "full")]


@pytest.mark.parametrize("n,incx,incy,kind", DOT_CASES)
def test_dot_exact_small(O, Q, cuda, n, incx, incy, kind):
    _say(f"dot exact n={n} incx={incx} incy={incy} kind={kind}")
    x, y = _vecs(O, kind, n, 11 + n, incx, incy)
    got = _dev_dot(Q, cuda, Q.MODE_EXACT, n, x, incx, y, incy)
    assert Q.dev_exact_status() == 0
    want, _ = O.exact_dot(x, y, incx=incx, incy=incy, n=n)
    assert got == want


def test_dot_exact_specials(O, Q, cuda):
    x, y = _vecs(O, "full", 300, 5)
    zero_x = x.copy()
    zero_x[:] = 0
    assert _dev_dot(Q, cuda, Q.MODE_EXACT, 300, zero_x, 1, y, 1) == (0, 0)
    # exact cancellation to +0
    x2 = np.concatenate([x[:10], x[:10]])
    y2 = np.concatenate([y[:10], y[:10]])
    x2[10:, 1] ^= np.uint64(1 << 63)
    assert _dev_dot(Q, cuda, Q.MODE_EXACT, 20, x2, 1, y2, 1) == (0, 0)
    # Inf and NaN follow the exact sum: one +inf product -> +inf; +inf and -inf products -> NaN
    inf = np.array(O.to_pair(O.q(float("inf"))), dtype=np.uint64)
    ya = O.abs_array(y)                               # all y_i > 0
    x3 = x.copy()
    x3[7] = inf
    r = _dev_dot(Q, cuda, Q.MODE_EXACT, 300, x3, 1, ya, 1)
    assert r == (0, 0x7FFF << 48)                     # +inf
    x3[8] = inf
    x3[8, 1] |= np.uint64(1 << 63)                    # -inf * y8 (> 0) = -inf
    r = _dev_dot(Q, cuda, Q.MODE_EXACT, 300, x3, 1, ya, 1)
    assert ((r[1] >> 48) & 0x7FFF) == 0x7FFF and ((r[1] & ((1 << 48) - 1)) | r[0]) != 0   # NaN
    # subnormal operands are exact too
    xs = O.fill(1, 3, 200, -16382, -16370)
    ys = O.fill(1, 4, 200, 5, 15)
    got = _dev_dot(Q, cuda, Q.MODE_EXACT, 200, xs, 1, ys, 1)
    want, _ = O.exact_dot(xs, ys)
    assert got == want


def test_dot_exact_full_size_and_properties(O, Q, cuda):
    n = 1 << 24
    for kind in ("full", "cancel"):
        _say(f"dot exact n=2^24 kind={kind}")
        x, y = _vecs(O, kind, n, 101)
        got = _dev_dot(Q, cuda, Q.MODE_EXACT, n, x, 1, y, 1)
        want, _ = O.exact_dot(x, y)
        assert got == want
        # correctly rounded => operand swap and permutation invariant
        assert _dev_dot(Q, cuda, Q.MODE_EXACT, n, y, 1, x, 1) == got
        perm = np.random.default_rng(3).permutation(n)
        assert _dev_dot(Q, cuda, Q.MODE_EXACT, n, np.ascontiguousarray(x[perm]), 1,
                        np.ascontiguousarray(y[perm]), 1) == got


def test_workspace_reuse_across_streams(O, Q, cuda):
    # a DD dot on one stream and a DD ColMajor gemv issued right after on another, both with
    # scratch memory from the library: each result equals its single-stream value bit for bit.
    # Then six streams in turn (more than the library keeps buffers for, so buffers are dropped
    # and reallocated between calls), each call checked the same way.
    n, m, k = 1 << 22, 2048, 2048
    x, y = _vecs(O, "full", n, 61)
    A, v, w0 = O.fill(0, 62, m * k), O.fill(0, 63, k), O.fill(0, 64, m)
    dx, dy, dA, dv = to_dev(x), to_dev(y), to_dev(A), to_dev(v)

    def run(s1, s2):
        dr, dw = dev_empty(1), to_dev(w0)
        cuda.cuda.synchronize()
        Q.dev_dot(Q.MODE_DD, n, dx.data_ptr(), 1, dy.data_ptr(), 1, 1, dr.data_ptr(), s1)
        Q.dev_gemv(Q.MODE_DD, 1, m, k, 1.5, dA.data_ptr(), m, dv.data_ptr(), 1, -0.5, dw.data_ptr(), 1, s2)
        cuda.cuda.synchronize()
        return from_dev(dr), from_dev(dw)

    ref_r, ref_w = run(0, 0)              # also grows the workspace to its final size
    sa, sb = cuda.cuda.Stream(), cuda.cuda.Stream()
    for s1, s2 in ((sa, sb), (sb, sa), (sa, sb)):
        _say("dd dot then dd gemv on two streams")
        r, w = run(s1.cuda_stream, s2.cuda_stream)
        assert np.array_equal(r, ref_r) and np.array_equal(w, ref_w)
    streams = [cuda.cuda.Stream() for _ in range(6)]
    for rnd in range(2):
        _say(f"dd dot then dd gemv over six streams, round {rnd}")
        for i in range(6):
            r, w = run(streams[i].cuda_stream, streams[(i + 1) % 6].cuda_stream)
            assert np.array_equal(r, ref_r) and np.array_equal(w, ref_w), (rnd, i)


GEMV_CASES = [(0, 1, 1, 1.0, 0.0), (0, 7, 300, 1.5, -0.5), (0, 64, 1000, "q", "q"), (0, 300, 33, 1.0, 1.0),
              (1, 1, 1, 1.0, 0.0), (1, 7, 300, 1.5, -0.5), (1, 513, 1000, "q", "q"), (1, 300, 33, 2.0, 0.0)]


@pytest.mark.parametrize("layout,m,n,alpha,beta", GEMV_CASES)
def test_gemv_exact_small(O, Q, cuda, layout, m, n, alpha, beta):
    _say(f"gemv exact layout={layout} m={m} n={n}")
    if alpha == "q":
        alpha, beta = O.fill(0, 41, 1)[0], O.fill(0, 42, 1)[0]
    lda = n if layout == 0 else m
    A = O.fill(0, 43 + m, m * n)
    x = O.fill(0, 44, n)
    y = O.fill(0, 45, m)
    want = y.copy()
    O.exact_gemv(layout, m, n, alpha, A, lda, x, 1, beta, want, 1)
    dA, dx, dy = to_dev(A), to_dev(x), to_dev(y)
    Q.dev_gemv(Q.MODE_EXACT, layout, m, n, alpha, dA.data_ptr(), lda, dx.data_ptr(), 1, beta, dy.data_ptr(), 1)
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() == 0
    assert np.array_equal(from_dev(dy), want)


def test_gemv_exact_full_size(O, Q, cuda):
    m = n = 8192
    A = O.fill(0, 3, m * n)
    x = O.fill(0, 4, n)
    y0 = O.fill(0, 5, m)
    want = y0.copy()
    O.exact_gemv(0, m, n, 1.5, A, n, x, 1, -0.5, want, 1)
    dA, dx = to_dev(A), to_dev(x)
    for layout in (0, 1):
        _say(f"gemv exact 8192x8192 layout={layout}")
        dy = to_dev(y0)
        # the same buffer read column-major is A^T: compare RowMajor(A) with ColMajor(A^T) data
        Q.dev_gemv(Q.MODE_EXACT, 0 if layout == 0 else 1, m, n, 1.5, dA.data_ptr(), n, dx.data_ptr(), 1, -0.5,
                   dy.data_ptr(), 1)
        cuda.cuda.synchronize()
        got = from_dev(dy)
        if layout == 0:
            assert np.array_equal(got, want)
        else:
            wantT = y0.copy()
            O.exact_gemv(1, m, n, 1.5, A, n, x, 1, -0.5, wantT, 1)
            assert np.array_equal(got, wantT)


def test_gemv_exact_row_equals_col_on_transposed_storage(O, Q, cuda):
    m, n = 300, 517
    A = O.fill(0, 7, m * n)                          # row-major m x n
    AT = np.ascontiguousarray(A.reshape(m, n, 2).transpose(1, 0, 2)).reshape(-1, 2)  # col-major m x n
    x = O.fill(0, 8, n)
    y = O.fill(0, 9, m)
    outs = []
    for layout, buf, lda in ((0, A, n), (1, AT, m)):
        dA, dx, dy = to_dev(buf), to_dev(x), to_dev(y)
        Q.dev_gemv(Q.MODE_EXACT, layout, m, n, 1.25, dA.data_ptr(), lda, dx.data_ptr(), 1, 0.75, dy.data_ptr(), 1)
        cuda.cuda.synchronize()
        outs.append(from_dev(dy))
    assert np.array_equal(outs[0], outs[1])


def test_exact_through_c_abi(O, Q, cuda):
    n = 5000
    x, y = _vecs(O, "cancel", n, 17)
    old = Q.get_mode()
    Q.set_mode("exact")
    try:
        want, _ = O.exact_dot(x, y)
        assert pair(Q.dot(n, x, 1, y, 1)) == want
        assert Q.qdot(n, x, 1, y, 1) == O.to_double(O.q(want))
        m, k = 40, 300
        A = O.fill(0, 61, m * k)
        xv = O.fill(0, 62, k)
        yv = O.fill(0, 63, m)
        want = yv.copy()
        O.exact_gemv(0, m, k, 1.5, A, k, xv, 1, -0.5, want, 1)
        got = yv.copy()
        Q.qgemv("R", "N", m, k, 1.5, A, k, xv, 1, -0.5, got, 1)
        assert np.array_equal(got, want)
    finally:
        Q.set_mode(old)


# ------------------------------------------------------------------------------------------------
# double-double form
# ------------------------------------------------------------------------------------------------

def _err_ok(got_pair, want_pair, scale):
    g, w = to_fraction(*got_pair), to_fraction(*want_pair)
    ulp = abs(w) * (2 ** -112) if w != 0 else 0
    return abs(g - w) <= DD_TOL * scale + ulp


@pytest.mark.parametrize("n,kind", [(1, "full"), (1000, "full"), (100003, "full"), (100003, "cancel"), (1 << 20, "full")])
def test_dot_dd_within_bound(O, Q, cuda, n, kind):
    _say(f"dot dd n={n} kind={kind}")
    x, y = _vecs(O, kind, n, 23)
    got = _dev_dot(Q, cuda, Q.MODE_DD, n, x, 1, y, 1)
    want, _ = O.exact_dot(x, y)
    scale, _ = O.exact_dot(O.abs_array(x), O.abs_array(y))
    assert _err_ok(got, want, to_fraction(*scale))


@pytest.mark.parametrize("layout,m,n", [(0, 64, 1000), (0, 1000, 64), (1, 64, 1000), (1, 1000, 64), (0, 999, 130),
                                        (1, 999, 130), (0, 1, 700), (0, 3, 257), (1, 3, 257), (0, 8192, 8192),
                                        (1, 8192, 8192)])
def test_gemv_dd_within_bound(O, Q, cuda, layout, m, n):
    _say(f"gemv dd layout={layout} m={m} n={n}")
    lda = n if layout == 0 else m
    A = O.fill(0, 51, m * n)
    x = O.fill(0, 52, n)
    y = O.fill(0, 53, m)
    want = y.copy()
    O.exact_gemv(layout, m, n, 1.5, A, lda, x, 1, -0.5, want, 1)
    scale = O.abs_array(y)
    O.exact_gemv(layout, m, n, 1.5, O.abs_array(A), lda, O.abs_array(x), 1, 0.5, scale, 1)
    dA, dx, dy = to_dev(A), to_dev(x), to_dev(y)
    Q.dev_gemv(Q.MODE_DD, layout, m, n, 1.5, dA.data_ptr(), lda, dx.data_ptr(), 1, -0.5, dy.data_ptr(), 1)
    cuda.cuda.synchronize()
    got = from_dev(dy)
    rows = range(m) if m <= 1000 else range(0, m, 97)
    for i in rows:
        assert _err_ok(pair(got[i]), pair(want[i]), to_fraction(*pair(scale[i]))), i


GEMM_DD_CASES = [(0, 64, 64, 64, "full"), (0, 100, 70, 300, "full"), (1, 100, 70, 300, "full"), (0, 65, 129, 33, "log"),
                 (1, 7, 300, 5, "log"), (0, 1, 1, 1, "full"), (0, 256, 192, 1000, "log")]


@pytest.mark.parametrize("layout,m,n,k,kind", GEMM_DD_CASES)
def test_gemm_dd_within_bound(O, Q, cuda, layout, m, n, k, kind):
    _say(f"gemm dd layout={layout} m={m} n={n} k={k} kind={kind}")
    lda, ldb, ldc = (k, n, n) if layout == 0 else (m, k, m)
    A, _ = _vecs(O, kind, m * k, 61 + m)
    B, _ = _vecs(O, kind, k * n, 62 + n)
    C0 = O.fill(0, 63, m * n)
    alpha, beta = 1.5, -0.5
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C0)
    Q.dev_gemm(Q.MODE_DD, layout, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc)
    cuda.cuda.synchronize()
    got = from_dev(dC)
    st = (lambda ld: (ld, 1)) if layout == 0 else (lambda ld: (1, ld))
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    want, _ = O.exact_gemm_samples(ii, jj, k, alpha, A, st(lda), B, st(ldb), beta, C0, st(ldc))
    scale, _ = O.exact_gemm_samples(ii, jj, k, alpha, O.abs_array(A), st(lda), O.abs_array(B), st(ldb), 0.5,
                                    O.abs_array(C0), st(ldc))
    cs = st(ldc)
    for s in range(ii.shape[0]):
        g = got[ii[s] * cs[0] + jj[s] * cs[1]]
        assert _err_ok(pair(g), pair(want[s]), to_fraction(*pair(scale[s]))), (ii[s], jj[s])


def test_global_mode_maps_per_routine(O, Q, cuda):
    # host-pointer entry points under one global mode: OZAKI has no dot, which then takes the EXACT
    # form; EXACT has no gemm, which then takes the OZAKI form (<= 2^-104 per entry)
    old = Q.get_mode()
    try:
        Q.set_mode("ozaki")
        x, y = _vecs(O, "cancel", 3000, 91)
        want, _ = O.exact_dot(x, y)
        assert pair(Q.dot(3000, x, 1, y, 1)) == want
        Q.set_mode("exact")
        m = n = k = 80
        A, B, C = O.fill(1, 92, m * k, -20, 20), O.fill(1, 93, k * n, -20, 20), O.fill(0, 94, m * n)
        got = C.copy()
        Q.gemm(0, m, n, k, 1.0, A, k, B, n, 0.5, got, n)
        ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        _, rel = O.exact_gemm_samples(ii.ravel(), jj.ravel(), k, 1.0, A, (k, 1), B, (n, 1), 0.5, C, (n, 1),
                                      cand=got)
        assert rel.max() <= 2.0 ** -104
    finally:
        Q.set_mode(old)
