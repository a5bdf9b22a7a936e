"""Pins the CPU oracle (oracle/qoracle.c) before anything is checked against it.

What pins what
  * arithmetic: or_fma bit-for-bit against libquadmath fmaq (six operand regimes, incl. exact ties,
    all-ones significands, subnormal/overflow edges, deep cancellation); add/mul/fma against exact
    rational arithmetic (Python Fractions) on random operands -- independent of libquadmath.
  * evaluation order of the restatements: the reference test suite's known answers
    (test_quadblas.cpp), exact integer closed forms, and hand-derived cases whose result differs
    between the reference's order and any other plausible order (2-lane even/odd chains, the odd
    tail, T-chunking with the remainder on the last chunk, kc = 126 restarted chains with beta on
    the first block only).  Derivations are written next to each case.
  * exact accumulator: against Fractions.
  * generators: against Fractions and determinism.
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

from helpers import pair, round_fraction, rows, to_fraction

B113 = 2 ** 113  # ulp(2^113) = 2, so 2^113 + 1 is an exact tie that rounds to 2^113 (even)


def qa(O, vals):
    """uint64 (n, 2) from Python ints/floats exactly representable in binary128 via doubles or
    as sums handled below."""
    out = np.zeros((len(vals), 2), dtype=np.uint64)
    for i, v in enumerate(vals):
        if isinstance(v, int) and abs(v) >= 2 ** 53:
            lo, hi = round_fraction(Fraction(v))
        else:
            r = O.q(float(v))
            lo, hi = r.lo, r.hi
        out[i] = (lo, hi)
    return out


def val(p):
    return to_fraction(*p)


# ------------------------------------------------------------------------------------------------
# arithmetic
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("mode", range(6))
def test_fma_matches_libquadmath(O, mode):
    bad, first = O.fma_selftest(1000 + mode, 200_000, mode)
    assert bad == 0, f"first mismatch (a, b, c, ours, fmaq): {first}"


def _rand_q(rng, emin=-40, emax=40):
    m = rng.getrandbits(112) | (1 << 112)
    e = rng.randint(emin, emax)
    s = rng.getrandbits(1)
    return round_fraction((-1) ** s * Fraction(m, 2 ** 112) * Fraction(2) ** e)


def test_arith_against_fractions(O):
    rng = random.Random(5)
    for i in range(3000):
        a, b = _rand_q(rng), _rand_q(rng)
        # addend near the product's magnitude half the time (cancellation)
        if i % 2:
            c = round_fraction(-(val(a) * val(b)) * (1 + Fraction(rng.randint(-1000, 1000), 2 ** 100)))
        else:
            c = _rand_q(rng, -90, 90)
        assert pair(np.array(O.to_pair(O.fma(a, b, c)), np.uint64)) == round_fraction(val(a) * val(b) + val(c))
        assert O.to_pair(O.add(a, c)) == round_fraction(val(a) + val(c))
        assert O.to_pair(O.mul(a, b)) == round_fraction(val(a) * val(b))


def test_fma_special_values(O):
    inf = O.q(float("inf"))
    nan = O.q(float("nan"))
    one, zero, mzero = O.q(1.0), O.q(0.0), O.q(-0.0)
    isnan = lambda r: ((r.hi >> 48) & 0x7FFF) == 0x7FFF and ((r.hi & ((1 << 48) - 1)) | r.lo) != 0
    assert isnan(O.fma(inf, zero, one))
    assert isnan(O.fma(inf, one, O.q(float("-inf"))))
    assert isnan(O.fma(nan, one, one))
    assert O.to_pair(O.fma(inf, one, one)) == O.to_pair(inf)
    assert O.to_pair(O.fma(zero, one, mzero)) == (0, 0)                 # +0 + -0 = +0
    assert O.to_pair(O.fma(mzero, one, mzero)) == (0, 1 << 63)          # -0 + -0 = -0
    assert O.to_pair(O.fma(one, one, O.q(-1.0))) == (0, 0)              # exact cancellation: +0


# ------------------------------------------------------------------------------------------------
# restatements: known answers from the reference's own tests
# ------------------------------------------------------------------------------------------------

def test_known_answer_dot_55(O):
    # test_quadblas.cpp:205-224: x = 1..10, y = 1 -> 55
    x = qa(O, range(1, 11))
    y = qa(O, [1] * 10)
    for T in (1, 2, 8):
        assert O.to_double(O.dot(10, x, 1, y, 1, T)) == 55.0


def test_known_answer_dot_15(O):
    # tests/test_sleef_simd.cpp and tests/debug_test.cpp: [1..5] . 1 = 15
    x = qa(O, range(1, 6))
    y = qa(O, [1] * 5)
    assert O.to_double(O.dot(5, x, 1, y, 1, 4)) == 15.0


def test_known_answer_cancellation(O):
    # test_quadblas.cpp:715-739: [1e20, 1, -1e20, 0 x7] . 1 = 1 exactly
    x = qa(O, [1e20, 1.0, -1e20] + [0.0] * 7)
    y = qa(O, [1.0] * 10)
    r = O.dot(10, x, 1, y, 1, 8)
    assert O.to_pair(r) == O.to_pair(O.q(1.0))


def test_axpy_known_answer_and_elementwise_fma(O):
    x = qa(O, [1, 2, 3])
    y = qa(O, [1, 1, 1])
    O.axpy(3, 2.0, x, 1, y, 1)
    assert [val(O.to_pair(O.q(p))) for p in y] == [3, 5, 7]
    # every element is one correctly rounded fma (level1.hpp:156, :207-220): compare with fmaq
    x = O.fill(0, 1, 50)
    y = O.fill(0, 2, 50)
    a = O.fill(0, 3, 1)[0]
    want = [O.to_pair(O.fmaq(a, x[i], y[i])) for i in range(50)]
    O.axpy(50, a, x, 1, y, 1)
    assert [tuple(int(v) for v in y[i]) for i in range(50)] == want


def test_known_answer_gemv(O):
    # test_quadblas.cpp:327-350: [[1,2,3],[4,5,6],[7,8,9]] . 1 = [6, 15, 24]
    A = qa(O, range(1, 10))
    x = qa(O, [1, 1, 1])
    y = qa(O, [0, 0, 0])
    O.gemv(0, 3, 3, 1.0, A, 3, x, 1, 0.0, y, 1)
    assert list(O.to_double(y)) == [6.0, 15.0, 24.0]
    # debug_test.cpp: [[1,2],[3,4]] . [1,1] = [3, 7]
    A = qa(O, [1, 2, 3, 4])
    x = qa(O, [1, 1])
    y = qa(O, [0, 0])
    O.gemv(0, 2, 2, 1.0, A, 2, x, 1, 0.0, y, 1)
    assert list(O.to_double(y)) == [3.0, 7.0]
    # the same matrix stored column-major gives the transposed product
    y = qa(O, [0, 0])
    O.gemv(1, 2, 2, 1.0, A, 2, x, 1, 0.0, y, 1)
    assert list(O.to_double(y)) == [4.0, 6.0]


def test_known_answer_gemm(O):
    # test_quadblas.cpp:446-464: [[1,2],[3,4]] [[5,6],[7,8]] = [[19,22],[43,50]]
    A = qa(O, [1, 2, 3, 4])
    B = qa(O, [5, 6, 7, 8])
    C = qa(O, [0, 0, 0, 0])
    O.gemm(0, 2, 2, 2, 1.0, A, 2, B, 2, 0.0, C, 2)
    assert list(O.to_double(C)) == [19.0, 22.0, 43.0, 50.0]
    # test_quadblas.cpp:679-688: 3 * 4 = 12
    A, B, C = qa(O, [3]), qa(O, [4]), qa(O, [0])
    O.gemm(0, 1, 1, 1, 1.0, A, 1, B, 1, 0.0, C, 1)
    assert O.to_double(C)[0] == 12.0


def test_known_answer_identity_and_patterns(O):
    # test_quadblas.cpp:691-712: I * A = A (full-mantissa A here), and :509-536's integer pattern
    n = 50
    I = qa(O, [1.0 if i == j else 0.0 for i in range(n) for j in range(n)])
    A = O.fill(0, 3, n * n)
    C = np.zeros((n * n, 2), np.uint64)
    O.gemm(0, n, n, n, 1.0, I, n, A, n, 0.0, C, n)
    assert np.array_equal(C, A)
    m = 64
    Ap = qa(O, [i + j + 1 for i in range(m) for j in range(m)])
    Bp = qa(O, [(i * j + 1) % 7 for i in range(m) for j in range(m)])
    Cp = np.zeros((m * m, 2), np.uint64)
    O.gemm(0, m, m, m, 1.0, Ap, m, Bp, m, 0.0, Cp, m)
    want = [sum((i + l + 1) * ((l * j + 1) % 7) for l in range(m)) for i in range(m) for j in range(m)]
    assert list(O.to_double(Cp)) == [float(w) for w in want]


def test_closed_forms(O):
    # sum i = n(n+1)/2 and sum i^2 = n(n+1)(2n+1)/6, exact in every order, every T, strided too
    n = 3001
    x = qa(O, range(1, n + 1))
    one = qa(O, [1] * n)
    for T in (1, 3, 7):
        assert O.to_double(O.dot(n, x, 1, one, 1, T)) == n * (n + 1) / 2
        assert O.to_double(O.dot(n, x, 1, x, 1, T)) == n * (n + 1) * (2 * n + 1) / 6
    x2 = qa(O, [v for i in range(1, n + 1) for v in (i, -7)])
    assert O.to_double(O.dot(n, x2, 2, x2, 2, 5)) == n * (n + 1) * (2 * n + 1) / 6


# ------------------------------------------------------------------------------------------------
# restatements: hand-derived order-sensitive cases
# ------------------------------------------------------------------------------------------------

def test_dot_two_lane_order(O):
    # level1.hpp:14-35.  x = [B, 1, -B, 1], y = 1 (B = 2^113):
    #   lane 0 (even i): fma(B,1,0) = B, fma(-B,1,B) = 0;  lane 1 (odd i): 1, 2;  add -> 2.
    #   A single chain gives B, B+1 -> B (tie to even), 0, 1  -> 1.
    x = qa(O, [B113, 1, -B113, 1])
    y = qa(O, [1] * 4)
    assert val(O.to_pair(O.dot(4, x, 1, y, 1, 8))) == 2
    # strided (incx = 2) uses one serial chain (level1.hpp:128-134): same data spread out -> 1
    xs = qa(O, [B113, 0, 1, 0, -B113, 0, 1, 0])
    assert val(O.to_pair(O.dot(4, xs, 2, y, 1, 8))) == 1


def test_dot_odd_tail(O):
    # level1.hpp:29-32.  x = [B, 1, -B, 1, 1]: lanes give 0 and 2, add -> 2, tail fma(1,1,2) = 3.
    # A single chain: B, B, 0, 1, 2 -> 2.
    x = qa(O, [B113, 1, -B113, 1, 1])
    y = qa(O, [1] * 5)
    assert val(O.to_pair(O.dot(5, x, 1, y, 1, 1))) == 3


def test_dot_thread_chunks(O):
    # level1.hpp:38-77 with n = 1000 >= 500.  x[0] = B, x[1] = 1, x[500] = -B, x[501] = 1, else 0.
    #   T = 1: one chunk; lane 0 sees B then -B (500 is even) -> 0; lane 1 sees 1, 1 -> 2;
    #          0 + 2 -> 2.
    #   T = 2: chunk 0 = [0,500): add(B, 1) = B (tie to even);  chunk 1 = [500,1000):
    #          add(-B, 1) = -(B-1) (exact);  (0 + B) + -(B-1) -> 1.
    #   T = 3: chunk = 333; x[500] is odd within chunk 1 (offset 167), x[501] even (168):
    #          chunk 1 = add(1, -B) = -(B-1); total again 1.
    #   T = 1000 (> n/2): chunk = 1, 999 one-element chunks use only the tail fma, chunk 999 too:
    #          sum from 0 in chunk order: B, B (tie), ..., -B at 500 -> 0, then 1 -> 1.
    n = 1000
    v = [0] * n
    v[0], v[1], v[500], v[501] = B113, 1, -B113, 1
    x = qa(O, v)
    y = qa(O, [1] * n)
    got = {T: val(O.to_pair(O.dot(n, x, 1, y, 1, T))) for T in (1, 2, 3, 1000)}
    assert got == {1: 2, 2: 1, 3: 1, 1000: 1}
    # The last chunk takes the remainder (level1.hpp:58-59): n = 1001, T = 2 gives chunks [0, 500)
    # and [500, 1001).  With 1 at offsets 500, 501 and B at 1000, chunk 1's lanes add to 2 and its
    # odd tail fma(B, 1, 2) = B + 2 is exact; chunk 0 is 0.
    v = [0] * 1001
    v[500], v[501], v[1000] = 1, 1, B113
    x = qa(O, v)
    y = qa(O, [1] * 1001)
    assert val(O.to_pair(O.dot(1001, x, 1, y, 1, 2))) == B113 + 2
    # The odd tail is applied after the lane sum (level1.hpp:27-32), which rounding distinguishes:
    v = [0] * 1001
    v[0], v[1], v[1000] = B113, 1, -B113
    x = qa(O, v)
    # T = 1: lanes: B and 1 -> add(B, 1) = B (tie); tail fma(-B, 1, B) = 0.  A tail-first order
    # would give B + (-B) = 0 then + 1 = 1.
    assert val(O.to_pair(O.dot(1001, x, 1, y, 1, 1))) == 0


def test_dot_strided_thread_chunks(O):
    # level1.hpp:93-120: a strided dot with n >= 500 cuts n into T chunks of n/T elements (the
    # last takes the remainder), each ONE serial FMA chain from +0, the partials added from +0 in
    # chunk order.  n = 1001, T = 3: chunks [0, 333), [333, 666), [666, 1001).  y = 1.
    #   chunk 0: x[0] = B                                   -> p0 = B
    #   chunk 1: B, 1, -B, 1 at 333..336: B, B+1 -> B (tie), 0, 1   -> p1 = 1
    #   chunk 2: 1 at 666 and 1 at 1000 (the remainder)     -> p2 = 2
    #   ((0 + B) + 1) -> B (tie), + 2 -> B + 2 (exact).
    # Each plausible slip gives another value: two lanes per chunk (p1 = 2) -> B + 4; one chain
    # over all n -> B; partials added in reverse -> (2 + 1) + B -> B + 4 (tie to even); the
    # remainder on the first chunk -> 2B + -(B - 2) + 1 -> B + 4; the remainder dropped -> B.
    n, T = 1001, 3
    v = [0] * n
    v[0], v[333], v[334], v[335], v[336], v[666], v[1000] = B113, B113, 1, -B113, 1, 1, 1
    xs = [0] * (2 * n)
    xs[::2] = v
    ys = [7] * (3 * n)          # values off the stride would show up
    ys[::3] = [1] * n
    x, y = qa(O, xs), qa(O, ys)
    assert val(O.to_pair(O.dot(n, x, 2, y, 3, T))) == B113 + 2
    # T = 1: one chain over all n: B, 2B, 2B + 1 -> 2B, B, B + 1 -> B (tie), ... -> B
    assert val(O.to_pair(O.dot(n, x, 2, y, 3, 1))) == B113


def test_gemv_row_vs_col_order(O):
    # level2.hpp:15-50 (row) and :53-82 (col) on the same mathematical A = [[B, 1, -B, 1]]:
    #   RowMajor: 2-lane chain over the row -> 2; y = fma(1, 2, mul(0, y)) = 2.
    #   ColMajor: y = mul(0,y) = 0; chain over columns: B, B+1 -> B, 0, 1 -> 1.
    A = qa(O, [B113, 1, -B113, 1])  # 1 x 4, both layouts coincide in memory (m = 1)
    x = qa(O, [1] * 4)
    y = qa(O, [0])
    O.gemv(0, 1, 4, 1.0, A, 4, x, 1, 0.0, y, 1)
    assert val(pair(y[0])) == 2
    y = qa(O, [0])
    O.gemv(1, 1, 4, 1.0, A, 1, x, 1, 0.0, y, 1)
    assert val(pair(y[0])) == 1
    # ColMajor pre-rounds alpha*x_j (level2.hpp:71): alpha = 1/3-ish full-mantissa value
    # against the row formula fma(alpha, sum, beta*y); checked on random data for difference
    rng_x = O.fill(0, 11, 64)
    A = O.fill(0, 12, 64 * 64)
    y1 = np.zeros((64, 2), np.uint64)
    y2 = np.zeros((64, 2), np.uint64)
    alpha = O.fill(0, 13, 1)[0]
    O.gemv(1, 64, 64, alpha, A, 64, rng_x, 1, 0.0, y1, 1)
    # explicit column formula
    ax = [O.mul(alpha, rng_x[j]) for j in range(64)]
    for i in range(64):
        acc = O.q(0.0)
        for j in range(64):
            acc = O.fma(A[j * 64 + i], ax[j], acc)
        y2[i] = O.to_pair(acc)
    assert np.array_equal(y1, y2)


def test_gemm_kc_restart_and_beta(O):
    # level3.hpp:244-300 (blocked, k > 64): kc = 126 (blocking.hpp:39-41, x86-64).
    # m = n = 1, k = 130: A = [B, 0 x125 | 1, 1, 0, 0], B-matrix all ones, C0 = 5, alpha 1, beta 2.
    #   block 0: t = B; C = fma(1, B, mul(2, 5)) = B + 10.
    #   block 1: t = 2; C = fma(1, 2, mul(1, B + 10)) = B + 12.
    # One chain over all 130 with C added last: B, then +1 twice -> B, B (ties) -> B + 10 = B + 10.
    assert O.blocking_kc(4096) == 126 and O.blocking_kc(100) == 100
    k = 130
    a = [0] * k
    a[0], a[126], a[127] = B113, 1, 1
    A = qa(O, a)
    Bm = qa(O, [1] * k)
    C = qa(O, [5])
    O.gemm(0, 1, 1, k, 1.0, A, k, Bm, 1, 2.0, C, 1)
    assert val(pair(C[0])) == B113 + 12
    # the same with k <= 64 and m = n = 1 is the simple path: one chain, beta applied once
    k = 60
    a = [0] * k
    a[0], a[1], a[2] = B113, 1, 1
    A = qa(O, a)
    Bm = qa(O, [1] * k)
    C = qa(O, [5])
    O.gemm(0, 1, 1, k, 1.0, A, k, Bm, 1, 2.0, C, 1)
    assert val(pair(C[0])) == B113 + 10


def test_gemm_layouts_agree_on_transposed_storage(O):
    # ColMajor gemm of (A, B) equals RowMajor gemm of the same buffers read as A^T, B^T with the
    # operands swapped: C^T = B^T A^T.  Pins the column-major indexing of the restatement.
    m, n, k = 37, 29, 131
    A = O.fill(0, 21, m * k)   # column-major m x k (lda = m)
    Bm = O.fill(0, 22, k * n)  # column-major k x n (ldb = k)
    C1 = O.fill(0, 23, m * n)
    C2 = C1.copy()
    O.gemm(1, m, n, k, 1.5, A, m, Bm, k, -0.5, C1, m)
    O.gemm(0, n, m, k, 1.5, Bm, k, A, m, -0.5, C2, m)
    assert np.array_equal(C1, C2)


# ------------------------------------------------------------------------------------------------
# exact accumulator and generators
# ------------------------------------------------------------------------------------------------

def test_exact_dot_against_fractions(O):
    for seed, kind in ((1, 0), (2, 1), (3, 2)):
        n = 200
        x = O.fill(kind, seed, n)
        y = O.fill(kind if kind != 2 else 0, seed + 100, n) if kind != 2 else O.fill_pair_y(seed, n)
        exact = sum(val(p) * val(q) for p, q in zip(rows(x), rows(y)))
        rounded, rel = O.exact_dot(x, y, cand=x[:1])
        assert rounded == round_fraction(exact)
        want = abs(val(pair(x[0])) - exact) / abs(exact)
        assert math.isclose(rel[0], float(want), rel_tol=1e-12)


def test_exact_dot_alpha_beta(O):
    x = O.fill(0, 5, 50)
    y = O.fill(0, 6, 50)
    al, be, c = O.fill(0, 7, 3)
    exact = val(pair(al)) * sum(val(p) * val(q) for p, q in zip(rows(x), rows(y))) + val(pair(be)) * val(pair(c))
    rounded, _ = O.exact_dot(x, y, alpha=al, beta=be, c=c)
    assert rounded == round_fraction(exact)


def test_cancellation_generators(O):
    # kinds 2 and 4: exact pairs +-(10^e u) on slots 0, 1 of every 8 (e = 20, 22), kind 0 elsewhere;
    # the matching y repeats its value on each pair, so x . y over a block of 8 is the 6 kind-0
    # products alone
    n = 64
    for kind, e in ((2, 20), (4, 22)):
        x = O.fill(kind, 3, n)
        y = O.fill_pair_y(3, n)
        k0 = O.fill(0, 3, n)
        for i, p in enumerate(rows(x)):
            if i % 8 < 2:
                u = 1 + Fraction(O.rand64(3, i - (i % 2)) >> 12, 2 ** 52)
                want = u * 10 ** e * (-1 if i % 2 else 1)
                assert p == round_fraction(want) and val(p) == want, (kind, i)
            else:
                assert p == pair(k0[i])
        assert np.array_equal(y[0::8], y[1::8])
        pairs = sum(val(pair(x[i])) * val(pair(y[i])) for i in range(n) if i % 8 < 2)
        assert pairs == 0


def test_generators(O):
    n = 64
    v = O.fill(0, 9, n)
    for i, p in enumerate(rows(v)):
        r0, r1 = O.rand64(9, 2 * i), O.rand64(9, 2 * i + 1)
        a = Fraction((r0 >> 11), 2 ** 52) - 1
        b = Fraction((r1 >> 11), 2 ** 52) - 1
        assert p == round_fraction(a + b * Fraction(1, 2 ** 60))
    v = O.fill(1, 9, n, -30, 30)
    for p in rows(v):
        e = ((p[1] >> 48) & 0x7FFF) - 16383
        assert -30 <= e < 30
    x = O.fill(2, 4, 64)
    y = O.fill_pair_y(4, 64)
    for p in range(0, 64, 8):
        assert val(pair(x[p])) == -val(pair(x[p + 1])) and abs(val(pair(x[p]))) >= Fraction(10) ** 20
        assert pair(y[p]) == pair(y[p + 1])
    assert np.array_equal(O.fill(0, 9, n), O.fill(0, 9, n))
