"""GPU parity of the OZAKI form of gemm, slice method (int8 slices on the matrix cores).

Contract: every entry of C = alpha*A*B + beta*C has relative error <= 2^-104 against the exact
value (the oracle's exact big-integer accumulation of the same entry), and entries the bound
cannot certify are recomputed exactly, so they equal the oracle's correctly rounded value bit for
bit.  Cases: small and ragged shapes in both layouts, k spanning several int32 accumulation passes,
forced repairs (few slices), cancellation, zero rows, Inf operands, alpha = 0, and the 4096^3
workload with log-uniform entries over [2^-30, 2^30] on 4096 sampled entries.
"""
import numpy as np
import pytest

from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu

TOL = 2.0 ** -104


def _say(msg):
    print(msg, flush=True)


def _strides(layout, ld):
    return (ld, 1) if layout == 0 else (1, ld)


def _data(O, kind, seed, count):
    if kind == "log10":
        return O.fill(1, seed, count, -10, 10)
    if kind == "log30":
        return O.fill(1, seed, count, -30, 30)
    return O.fill(0, seed, count)


def _run(Q, cuda, layout, m, n, k, alpha, beta, A, lda, B, ldb, C, ldc, S):
    Q.set_ozaki_slices(S)
    Q.set_ozaki_method(Q.OZAKI_SLICES)
    try:
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        Q.dev_gemm(Q.MODE_OZAKI, layout, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(),
                   ldc)
        cuda.cuda.synchronize()
        status = Q.dev_exact_status()
        repairs = Q.dev_ozaki_repairs()
    finally:
        Q.set_ozaki_slices(0)
        Q.set_ozaki_method(Q.OZAKI_AUTO)
    return from_dev(dC), status, repairs


def _exact(O, layout, k, alpha, A, lda, B, ldb, beta, C0, ldc, got, ii, jj):
    cs = _strides(layout, ldc)
    cand = np.ascontiguousarray(got[ii * cs[0] + jj * cs[1]])
    rounded, rel = O.exact_gemm_samples(ii, jj, k, alpha, A, _strides(layout, lda), B, _strides(layout, ldb), beta,
                                        C0, cs, cand=cand)
    return rounded, rel, cand


def _all_idx(m, n):
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    return ii.ravel(), jj.ravel()


def _operands(O, layout, m, n, k, kind, seed):
    lda = k if layout == 0 else m
    ldb = n if layout == 0 else k
    ldc = n if layout == 0 else m
    A = _data(O, kind, seed, m * k)
    B = _data(O, kind, seed + 1, k * n)
    C = _data(O, "full", seed + 2, m * n)
    return A, lda, B, ldb, C, ldc


CASES = [
    (0, 64, 64, 64, "log10", 1.0, 0.0, 16),
    (0, 1, 1, 1, "full", 1.0, 0.0, 17),
    (0, 100, 70, 300, "full", 1.5, -0.5, 17),
    (1, 100, 70, 300, "full", 1.5, -0.5, 17),
    (0, 65, 129, 33, "log30", "q", "q", 17),
    (1, 7, 300, 5, "log30", 1.0, 1.0, 18),
    (0, 96, 64, 4100, "log30", 1.0, 0.0, 18),     # two int32 accumulation passes
    (1, 64, 96, 9000, "log10", 2.0, 0.25, 16),    # three passes, ColMajor
    (0, 128, 192, 512, "log30", 1.0, 0.0, 12),    # few slices: mostly repaired
    (0, 96, 64, 300, "log30", 1.0, 0.0, 20),      # 20..28 slices: two diagonal ranges
    (1, 100, 70, 4200, "full", 1.5, -0.5, 24),    # two ranges x two k passes, ColMajor
    (0, 65, 129, 33, "log30", "q", "q", 28),
]


@pytest.mark.parametrize("layout,m,n,k,kind,alpha,beta,S", CASES)
def test_ozaki_all_entries(O, Q, cuda, layout, m, n, k, kind, alpha, beta, S):
    _say(f"ozaki layout={layout} m={m} n={n} k={k} kind={kind} S={S}")
    if alpha == "q":
        alpha, beta = O.fill(0, 71, 1)[0], O.fill(0, 72, 1)[0]
    A, lda, B, ldb, C0, ldc = _operands(O, layout, m, n, k, kind, 1000 + m + n + k)
    got, status, repairs = _run(Q, cuda, layout, m, n, k, alpha, beta, A, lda, B, ldb, C0, ldc, S)
    assert status == 0
    assert 0 <= repairs <= m * n
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, layout, k, alpha, A, lda, B, ldb, beta, C0, ldc, got, ii, jj)
    _say(f"  repairs={repairs} max rel={rel.max():.3e} exact-equal={np.mean(np.all(cand == rounded, axis=1)):.3f}")
    assert rel.max() <= TOL


def test_ozaki_few_slices_repairs_are_correctly_rounded(O, Q, cuda):
    # with S = 8 the bound certifies almost nothing: nearly every entry is recomputed exactly
    m, n, k = 64, 64, 200
    A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, "full", 5)
    got, status, repairs = _run(Q, cuda, 0, m, n, k, 1.25, -0.75, A, lda, B, ldb, C0, ldc, 8)
    assert status == 0
    assert repairs >= m * n // 2
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, 0, k, 1.25, A, lda, B, ldb, -0.75, C0, ldc, got, ii, jj)
    assert rel.max() <= TOL
    # a repaired entry is the correctly rounded value; certified ones are within the tolerance
    assert np.mean(np.all(cand == rounded, axis=1)) >= repairs / (m * n) - 1e-9


def test_ozaki_cancellation_rows(O, Q, cuda):
    # rows of A made of +-large pairs (exact cancellation against equal B rows) plus a small residue:
    # the entries' magnitudes are far below the row/column scales, so they are repaired
    m, n, k = 64, 64, 256
    A = O.fill(0, 21, m * k)
    B = O.fill(0, 22, k * n)
    big = O.fill(1, 23, m * (k // 4), 60, 70).reshape(m, k // 4, 2)
    A3 = A.reshape(m, k, 2)
    A3[:, 0:k // 2:2] = big
    A3[:, 1:k // 2:2] = big
    A3[:, 1:k // 2:2, 1] ^= np.uint64(1 << 63)
    B3 = B.reshape(k, n, 2)
    B3[1:k // 2:2] = B3[0:k // 2:2]
    A = A3.reshape(-1, 2)
    B = B3.reshape(-1, 2)
    C0 = O.fill(0, 24, m * n)
    got, status, repairs = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A, k, B, n, C0, n, 17)
    assert status == 0
    assert repairs > 0
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, 0, k, 1.0, A, k, B, n, 0.0, C0, n, got, ii, jj)
    assert rel.max() <= TOL


def test_ozaki_zero_rows_and_alpha_zero(O, Q, cuda):
    m, n, k = 70, 40, 90
    A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, "log10", 31)
    A.reshape(m, k, 2)[5] = 0                     # zero row of A
    B.reshape(k, n, 2)[:, 7] = 0                  # zero column of B
    got, status, _ = _run(Q, cuda, 0, m, n, k, 1.5, -0.5, A, lda, B, ldb, C0, ldc, 17)
    assert status == 0
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, 0, k, 1.5, A, lda, B, ldb, -0.5, C0, ldc, got, ii, jj)
    assert rel.max() <= TOL
    zero_entries = (ii == 5) | (jj == 7)
    assert np.array_equal(cand[zero_entries], rounded[zero_entries])
    # alpha = 0: C = RN(beta * C)
    got, status, _ = _run(Q, cuda, 0, m, n, k, 0.0, 0.375, A, lda, B, ldb, C0, ldc, 17)
    rounded, rel, cand = _exact(O, 0, k, 0.0, A, lda, B, ldb, 0.375, C0, ldc, got, ii, jj)
    assert np.array_equal(cand, rounded)


def test_ozaki_inf_operand(O, Q, cuda):
    m, n, k = 40, 30, 50
    A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, "log10", 41)
    B = O.abs_array(B)                            # all B > 0
    inf = np.array(O.to_pair(O.q(float("inf"))), dtype=np.uint64)
    A.reshape(m, k, 2)[3, 9] = inf                # +inf in row 3
    got, status, repairs = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A, lda, B, ldb, C0, ldc, 17)
    assert repairs >= n
    g = got.reshape(m, n, 2)
    assert all(int(g[3, j, 1]) == (0x7FFF << 48) and int(g[3, j, 0]) == 0 for j in range(n))   # +inf
    # every other row is finite and within tolerance
    rows = np.array([i for i in range(m) if i != 3])
    ii, jj = np.meshgrid(rows, np.arange(n), indexing="ij")
    A_f = A.copy()
    A_f.reshape(m, k, 2)[3] = 0
    rounded, rel, cand = _exact(O, 0, k, 1.0, A_f, lda, B, ldb, 0.0, C0, ldc, got, ii.ravel(), jj.ravel())
    assert rel.max() <= TOL


def test_ozaki_4096_cube_sampled(O, Q, cuda):
    N = 4096
    _say("ozaki 4096^3 log-uniform [2^-30, 2^30], 4096 sampled entries")
    A = O.fill(1, 11, N * N, -30, 30)
    B = O.fill(1, 12, N * N, -30, 30)
    C0 = np.zeros((N * N, 2), dtype=np.uint64)
    # automatic slice count (16-17 here), which also takes the repair through a transposed copy of B
    # when the queue is long
    got, status, repairs = _run(Q, cuda, 0, N, N, N, 1.0, 0.0, A, N, B, N, C0, N, 0)
    assert status == 0
    assert 16 <= Q.ozaki_last_slices() <= 18
    rng = np.random.default_rng(2024)
    ii = rng.integers(0, N, 4096)
    jj = rng.integers(0, N, 4096)
    rounded, rel, cand = _exact(O, 0, N, 1.0, A, N, B, N, 0.0, C0, N, got, ii, jj)
    _say(f"  repairs={repairs} of {N * N}; max rel {rel.max():.3e} (2^{np.log2(max(rel.max(), 1e-300)):.1f}), "
         f"median 2^{np.log2(max(np.median(rel), 1e-300)):.1f}")
    assert rel.max() <= TOL


def test_fill_2d_matches_host_layout(O, Q, cuda):
    m, k = 24, 40
    d = cuda.empty((m * k, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_fill_2d(2, 9, m, k, m, 1, d.data_ptr())          # ColMajor m x k, value r*k + c
    cuda.cuda.synchronize()
    want = O.fill(2, 9, m * k).reshape(m, k, 2).transpose(1, 0, 2).reshape(-1, 2)
    assert np.array_equal(from_dev(d), want)


def test_ozaki_ill_conditioned_colmajor(O, Q, cuda):
    # ColMajor 64 x 64 x 16384: A rows carry +-1e20 u pairs along k, B columns equal values on
    # each pair, so the products of each pair cancel exactly and the entries are ~1e-18 of the
    # largest products (|C| ~ 2^-68 of 2^(E_i+F_j)).  26 slices certify all but entries within
    # ~2^-12 of zero; 17 slices certify nothing and every entry is recomputed exactly.
    m, n, k = 64, 64, 16384
    dA = cuda.empty((m * k, 2), dtype=cuda.int64, device="cuda:0")
    dB = cuda.empty((k * n, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_fill_2d(2, 5, m, k, m, 1, dA.data_ptr())         # A(i, l) = value i*k + l, at i + l*m
    Q.dev_fill_2d(3, 5, n, k, k, 0, dB.data_ptr())         # B(l, j) = value j*k + l, at l + j*k
    cuda.cuda.synchronize()
    A, B = from_dev(dA), from_dev(dB)
    C0 = np.zeros((m * n, 2), dtype=np.uint64)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    for S, max_repairs in ((26, 4), (17, None)):
        _say(f"ozaki ill-conditioned ColMajor 64x64x16384 S={S}")
        got, status, repairs = _run(Q, cuda, 1, m, n, k, 1.0, 0.0, A, m, B, k, C0, m, S)
        assert status == 0
        rounded, rel, cand = _exact(O, 1, k, 1.0, A, m, B, k, 0.0, C0, m, got, ii, jj)
        big = abs(O.to_double(O.q(tuple(int(v) for v in A[0]))))
        _say(f"  repairs={repairs}; max rel 2^{np.log2(max(rel.max(), 1e-300)):.1f}; |C| ~ "
             f"{np.median(np.abs(O.to_double(rounded))):.3g} vs |A| up to {big:.3g}")
        if max_repairs is not None:
            assert repairs <= max_repairs
        else:
            assert repairs == m * n
        assert rel.max() <= TOL


def test_ozaki_automatic_slice_count(O, Q, cuda):
    # the library's choice (slice count 0): a probe of 256 exact entries sets S so the bound
    # certifies all but the two lowest of them: 113-bit mantissas at k = 700 need 16-17 slices
    # whether the entries span one or sixty binades; the ill-conditioned workload (entries ~2^-68
    # below their scale) needs about 26 -- and then about 1% or less is left to the exact repair
    chosen = {}
    m = n = 192
    k = 700
    for kind in ("full", "log30"):
        A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, kind, 81)
        got, status, repairs = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A, lda, B, ldb, C0, ldc, 0)
        S = Q.ozaki_last_slices()
        ii, jj = _all_idx(m, n)
        _, rel, _ = _exact(O, 0, k, 1.0, A, lda, B, ldb, 0.0, C0, ldc, got, ii, jj)
        _say(f"auto slices {kind}: S={S} repairs={repairs} max rel 2^{np.log2(max(rel.max(), 1e-300)):.1f}")
        assert status == 0 and rel.max() <= TOL
        assert repairs <= m * n // 50
        chosen[kind] = S
    m, n, k = 64, 64, 16384
    dA = cuda.empty((m * k, 2), dtype=cuda.int64, device="cuda:0")
    dB = cuda.empty((k * n, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_fill_2d(2, 5, m, k, m, 1, dA.data_ptr())
    Q.dev_fill_2d(3, 5, n, k, k, 0, dB.data_ptr())
    cuda.cuda.synchronize()
    A, B = from_dev(dA), from_dev(dB)
    C0 = np.zeros((m * n, 2), dtype=np.uint64)
    got, status, repairs = _run(Q, cuda, 1, m, n, k, 1.0, 0.0, A, m, B, k, C0, m, 0)
    chosen["ill"] = Q.ozaki_last_slices()
    ii, jj = _all_idx(m, n)
    _, rel, _ = _exact(O, 1, k, 1.0, A, m, B, k, 0.0, C0, m, got, ii, jj)
    _say(f"auto slices ill-conditioned: S={chosen['ill']} repairs={repairs}")
    assert status == 0 and rel.max() <= TOL and repairs <= 40
    assert 15 <= chosen["full"] <= chosen["log30"] <= 19 and chosen["ill"] >= 24
