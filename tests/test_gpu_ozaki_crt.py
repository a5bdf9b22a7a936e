"""GPU parity of the OZAKI form of gemm, CRT method (residues modulo pairwise coprime moduli <= 256,
one int8 GEMM per modulus on the matrix cores, Chinese remainder reconstruction).

Same contract as the slice method: every entry of C = alpha*A*B + beta*C has relative error
<= 2^-104 against the exact value (the oracle's exact big-integer accumulation of the same entry),
and entries the bound cannot certify are recomputed exactly.  The reconstruction is exact, so the
certified value is the correctly rounded alpha*X*2^(E_i+F_j-2t) + beta*c of the exact product X of
the scaled integers; the tolerance covers the operands' rounding to t bits.  Cases: small and ragged
shapes in both layouts, more than one 256 x 256 product tile, k long enough for the in-loop modular
reduction (k > 65536), the fewest and most bits (13 and 46 moduli), zero rows, Inf operands,
alpha = 0, the automatic choice of method and bits, more tiles than workgroups of the persistent
product kernel with a short last stage of k blocks, and the 4096^3 workload on 4096 sampled
entries.
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


def _run(Q, cuda, layout, m, n, k, alpha, beta, A, lda, B, ldb, C, ldc, t, method=None):
    Q.set_ozaki_method(Q.OZAKI_CRT if method is None else method)
    Q.set_ozaki_crt_bits(t)
    try:
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
        Q.dev_gemm(Q.MODE_OZAKI, layout, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(),
                   ldc)
        cuda.cuda.synchronize()
        status = Q.dev_exact_status()
        repairs = Q.dev_ozaki_repairs()
        info = (Q.ozaki_last_method(),) + Q.ozaki_last_crt()
    finally:
        Q.set_ozaki_crt_bits(0)
        Q.set_ozaki_method(Q.OZAKI_AUTO)
    return from_dev(dC), status, repairs, info


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


def _moduli(k, t):
    # the fewest moduli (largest first) whose product is >= k 2^(2t-2)
    mods = [256, 251, 243, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191, 181, 179, 173, 169, 167, 163,
            157, 151, 149, 139, 137, 131, 127, 125, 121, 113, 109, 107, 103, 101, 97, 89, 83, 79, 73, 71, 67, 61,
            59, 53, 49, 47, 43]
    need = 1 << (2 * t - 2 + max(0, (k - 1).bit_length()))
    prod = 1
    for q, p in enumerate(mods):
        prod *= p
        if prod >= need:
            return q + 1
    return 0


CASES = [
    (0, 64, 64, 64, "log10", 1.0, 0.0, 130),
    (0, 1, 1, 1, "full", 1.0, 0.0, 136),
    (0, 100, 70, 300, "full", 1.5, -0.5, 136),
    (1, 100, 70, 300, "full", 1.5, -0.5, 136),
    (0, 65, 129, 33, "log30", "q", "q", 140),
    (1, 7, 300, 5, "log30", 1.0, 1.0, 144),
    (0, 300, 257, 70, "log30", 1.0, 0.0, 150),     # 2 x 2 product tiles
    (1, 260, 40, 130, "log10", -1.0, 2.0, 138),    # ColMajor, two tile rows
    (0, 96, 64, 4100, "log30", 1.0, 0.0, 146),
    (1, 64, 96, 9000, "log10", 2.0, 0.25, 130),
    (0, 128, 192, 512, "log30", 1.0, 0.0, 60),     # few bits: mostly repaired
    (0, 65, 129, 33, "log30", "q", "q", 158),      # the most bits
    (1, 33, 40, 96, "full", 1.0, 0.0, 48),         # the fewest bits
]


@pytest.mark.parametrize("layout,m,n,k,kind,alpha,beta,t", CASES)
def test_crt_all_entries(O, Q, cuda, layout, m, n, k, kind, alpha, beta, t):
    _say(f"ozaki crt layout={layout} m={m} n={n} k={k} kind={kind} t={t}")
    if alpha == "q":
        alpha, beta = O.fill(0, 71, 1)[0], O.fill(0, 72, 1)[0]
    A, lda, B, ldb, C0, ldc = _operands(O, layout, m, n, k, kind, 2000 + m + n + k)
    got, status, repairs, info = _run(Q, cuda, layout, m, n, k, alpha, beta, A, lda, B, ldb, C0, ldc, t)
    assert info == (Q.OZAKI_CRT, t, _moduli(k, t))
    assert status == 0
    assert 0 <= repairs <= m * n
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, layout, k, alpha, A, lda, B, ldb, beta, C0, ldc, got, ii, jj)
    _say(f"  moduli={info[2]} repairs={repairs} max rel={rel.max():.3e} "
         f"exact-equal={np.mean(np.all(cand == rounded, axis=1)):.3f}")
    assert rel.max() <= TOL


def test_crt_well_conditioned_needs_no_repair(O, Q, cuda):
    # full-mantissa uniform entries at t = 140: the bound certifies every entry
    m, n, k = 128, 96, 500
    A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, "full", 7)
    got, status, repairs, info = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A, lda, B, ldb, C0, ldc, 140)
    assert status == 0 and repairs == 0
    ii, jj = _all_idx(m, n)
    _, rel, _ = _exact(O, 0, k, 1.0, A, lda, B, ldb, 0.0, C0, ldc, got, ii, jj)
    assert rel.max() <= TOL


def test_crt_long_k_modular_reduction(O, Q, cuda):
    # k = 65600: the product accumulators are reduced modulo p after 65536 terms
    m, n, k = 40, 36, 65600
    A, lda, B, ldb, C0, ldc = _operands(O, 1, m, n, k, "log10", 9)
    got, status, repairs, info = _run(Q, cuda, 1, m, n, k, 1.0, -1.0, A, lda, B, ldb, C0, ldc, 140)
    assert status == 0
    rng = np.random.default_rng(5)
    ii, jj = rng.integers(0, m, 96), rng.integers(0, n, 96)
    _, rel, _ = _exact(O, 1, k, 1.0, A, lda, B, ldb, -1.0, C0, ldc, got, ii, jj)
    _say(f"  k=65600 moduli={info[2]} repairs={repairs} max rel={rel.max():.3e}")
    assert rel.max() <= TOL


def test_crt_zero_rows_alpha_zero_and_inf(O, Q, cuda):
    m, n, k = 70, 40, 90
    A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, "log10", 31)
    A.reshape(m, k, 2)[5] = 0
    B.reshape(k, n, 2)[:, 7] = 0
    got, status, _, _ = _run(Q, cuda, 0, m, n, k, 1.5, -0.5, A, lda, B, ldb, C0, ldc, 136)
    assert status == 0
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, 0, k, 1.5, A, lda, B, ldb, -0.5, C0, ldc, got, ii, jj)
    assert rel.max() <= TOL
    zero_entries = (ii == 5) | (jj == 7)
    assert np.array_equal(cand[zero_entries], rounded[zero_entries])
    got, status, _, _ = _run(Q, cuda, 0, m, n, k, 0.0, 0.375, A, lda, B, ldb, C0, ldc, 136)
    rounded, rel, cand = _exact(O, 0, k, 0.0, A, lda, B, ldb, 0.375, C0, ldc, got, ii, jj)
    assert np.array_equal(cand, rounded)
    # +inf in row 3 of A (B > 0): row 3 is +inf, the rest within tolerance
    B = O.abs_array(B)
    inf = np.array(O.to_pair(O.q(float("inf"))), dtype=np.uint64)
    A2 = A.copy()
    A2.reshape(m, k, 2)[3, 9] = inf
    got, status, repairs, _ = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A2, lda, B, ldb, C0, ldc, 136)
    g = got.reshape(m, n, 2)
    assert all(int(g[3, j, 1]) == (0x7FFF << 48) and int(g[3, j, 0]) == 0 for j in range(n) if j != 7)
    rows = np.array([i for i in range(m) if i != 3])
    ri, rj = np.meshgrid(rows, np.arange(n), indexing="ij")
    A_f = A2.copy()
    A_f.reshape(m, k, 2)[3] = 0
    _, rel, _ = _exact(O, 0, k, 1.0, A_f, lda, B, ldb, 0.0, C0, ldc, got, ri.ravel(), rj.ravel())
    assert rel.max() <= TOL


def test_crt_automatic_choice(O, Q, cuda):
    # AUTO: the probe sets the bits; a 512 x 512 x 700 product with log-uniform entries takes the
    # CRT method, the ill-conditioned workload (entries ~2^-68 below their scale) needs more than
    # 158 bits and takes the slice method
    m = n = 512
    k = 700
    A, lda, B, ldb, C0, ldc = _operands(O, 0, m, n, k, "log30", 81)
    got, status, repairs, info = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A, lda, B, ldb, C0, ldc, 0, Q.OZAKI_AUTO)
    _say(f"auto log30 512^2 x 700: method={info[0]} bits={info[1]} moduli={info[2]} repairs={repairs}")
    assert status == 0 and info[0] == Q.OZAKI_CRT and 110 <= info[1] <= 150
    assert repairs <= m * n // 50
    rng = np.random.default_rng(1)
    ii, jj = rng.integers(0, m, 2048), rng.integers(0, n, 2048)
    _, rel, _ = _exact(O, 0, k, 1.0, A, lda, B, ldb, 0.0, C0, ldc, got, ii, jj)
    assert rel.max() <= TOL
    m, n, k = 64, 64, 16384
    dA = cuda.empty((m * k, 2), dtype=cuda.int64, device="cuda:0")
    dB = cuda.empty((k * n, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_fill_2d(2, 5, m, k, m, 1, dA.data_ptr())
    Q.dev_fill_2d(3, 5, n, k, k, 0, dB.data_ptr())
    cuda.cuda.synchronize()
    A, B = from_dev(dA), from_dev(dB)
    C0 = np.zeros((m * n, 2), dtype=np.uint64)
    got, status, repairs, info = _run(Q, cuda, 1, m, n, k, 1.0, 0.0, A, m, B, k, C0, m, 0, Q.OZAKI_AUTO)
    _say(f"auto ill-conditioned: method={info[0]} slices={Q.ozaki_last_slices()} repairs={repairs}")
    assert status == 0 and info[0] == Q.OZAKI_SLICES and Q.ozaki_last_slices() >= 24


def test_crt_4096_cube_sampled(O, Q, cuda):
    N = 4096
    _say("ozaki crt 4096^3 log-uniform [2^-30, 2^30], automatic method and bits, 4096 sampled entries")
    A = O.fill(1, 11, N * N, -30, 30)
    B = O.fill(1, 12, N * N, -30, 30)
    C0 = np.zeros((N * N, 2), dtype=np.uint64)
    got, status, repairs, info = _run(Q, cuda, 0, N, N, N, 1.0, 0.0, A, N, B, N, C0, N, 0, Q.OZAKI_AUTO)
    _say(f"  method={info[0]} bits={info[1]} moduli={info[2]} repairs={repairs}")
    assert status == 0 and info[0] == Q.OZAKI_CRT
    rng = np.random.default_rng(2024)
    ii = rng.integers(0, N, 4096)
    jj = rng.integers(0, N, 4096)
    rounded, rel, cand = _exact(O, 0, N, 1.0, A, N, B, N, 0.0, C0, N, got, ii, jj)
    _say(f"  max rel {rel.max():.3e} (2^{np.log2(max(rel.max(), 1e-300)):.1f}), "
         f"median 2^{np.log2(max(np.median(rel), 1e-300)):.1f}")
    assert rel.max() <= TOL


@pytest.mark.parametrize("layout,k", [(0, 4100), (1, 4160), (0, 4192)])
def test_crt_tiles_per_workgroup_short_last_stage(O, Q, cuda, layout, k):
    # 1024 x 1024 at t = 136: 16 tiles x about 40 moduli, more than one tile per workgroup of the
    # persistent product kernel (one workgroup per CU), so each workgroup copies its next tile's
    # first stage during the last stage of the current one; k = 4100, 4160, 4192 leave 1, 2, 3 k
    # blocks in the short last stage of four
    m = n = 1024
    t = 136
    _say(f"ozaki crt persistent tiles layout={layout} m=n={m} k={k} t={t}")
    A, lda, B, ldb, C0, ldc = _operands(O, layout, m, n, k, "log30", 3100 + k)
    got, status, repairs, info = _run(Q, cuda, layout, m, n, k, 1.5, -0.5, A, lda, B, ldb, C0, ldc, t)
    assert info == (Q.OZAKI_CRT, t, _moduli(k, t)) and 16 * info[2] > 256
    assert status == 0
    rng = np.random.default_rng(k)
    ii = np.concatenate([rng.integers(0, m, 700), np.full(16, m - 1), np.arange(16)])
    jj = np.concatenate([rng.integers(0, n, 700), np.arange(n - 16, n), np.full(16, n - 1)])
    rounded, rel, cand = _exact(O, layout, k, 1.5, A, lda, B, ldb, -0.5, C0, ldc, got, ii, jj)
    _say(f"  moduli={info[2]} repairs={repairs} max rel={rel.max():.3e}")
    assert rel.max() <= TOL


@pytest.mark.parametrize("method", ["auto", "slices", "crt"])
def test_repair_reads_b_through_transposed_copy(O, Q, cuda, method):
    # RowMajor A(i, 2l) = A(i, 2l+1), B(2l+1, j) = -B(2l, j): every entry cancels exactly to zero.
    # With beta = 0 the result is the product alone, which no bound certifies (the CRT method's
    # exact product is zero; the slice method's is its truncation error), so all 9216 entries are
    # queued; with the probe run (the call synchronizes) a queue of >= 2048 entries makes the
    # repair read B's columns from a transposed copy.  Every entry must equal the oracle's value.
    m, n, k = 96, 96, 64
    half = O.fill(0, 91, m * (k // 2)).reshape(m, k // 2, 2)
    A = np.repeat(half, 2, axis=1).reshape(-1, 2)
    bh = O.fill(0, 92, (k // 2) * n).reshape(k // 2, n, 2)
    B3 = np.repeat(bh, 2, axis=0)
    B3[1::2, :, 1] ^= np.uint64(1 << 63)
    B = B3.reshape(-1, 2)
    C0 = O.fill(0, 93, m * n)
    meth = {"auto": Q.OZAKI_AUTO, "slices": Q.OZAKI_SLICES, "crt": Q.OZAKI_CRT}[method]
    got, status, repairs, info = _run(Q, cuda, 0, m, n, k, 1.0, 0.0, A, k, B, n, C0, n, 0, meth)
    assert status == 0 and repairs == m * n
    ii, jj = _all_idx(m, n)
    rounded, rel, cand = _exact(O, 0, k, 1.0, A, k, B, n, 0.0, C0, n, got, ii, jj)
    assert np.array_equal(cand, rounded)


def test_ill_conditioned_full_size_forms(O, Q, cuda):
    # The ill-conditioned workload at full size: qgemm 8192 x 8192 x 16384 ColMajor, alpha = 1,
    # beta = 0, entries ~1e-20 of the largest products (+-1e22 u pairs, generator kind 4).  The
    # OZAKI form (automatic: the probe wants more than the CRT method's 158 bits, so the slice
    # method) meets 2^-104 on 128 sampled entries; the DD and plain forms are run on the same data
    # and their errors against the exact value recorded.
    m, n, k = 8192, 8192, 16384
    dA = cuda.empty((m * k, 2), dtype=cuda.int64, device="cuda:0")
    dB = cuda.empty((k * n, 2), dtype=cuda.int64, device="cuda:0")
    dC = cuda.zeros((m * n, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_fill_2d(4, 5, m, k, m, 1, dA.data_ptr())
    Q.dev_fill_2d(3, 5, n, k, k, 0, dB.data_ptr())
    cuda.cuda.synchronize()
    A, B = from_dev(dA), from_dev(dB)
    C0 = np.zeros((m * n, 2), dtype=np.uint64)
    rng = np.random.default_rng(8)
    ii, jj = rng.integers(0, m, 128), rng.integers(0, n, 128)
    worst = {}
    for form, mode in (("ozaki", Q.MODE_OZAKI), ("dd", Q.MODE_DD), ("plain", Q.MODE_REFERENCE)):
        _say(f"ill-conditioned 8192x8192x16384 ColMajor, {form}")
        dC.zero_()
        Q.dev_gemm(mode, 1, m, n, k, 1.0, dA.data_ptr(), m, dB.data_ptr(), k, 0.0, dC.data_ptr(), m)
        cuda.cuda.synchronize()
        if form == "ozaki":
            assert Q.dev_exact_status() == 0
            assert Q.ozaki_last_method() == Q.OZAKI_SLICES
        _, rel, _ = _exact(O, 1, k, 1.0, A, m, B, k, 0.0, C0, m, from_dev(dC), ii, jj)
        worst[form] = float(rel.max())
        _say(f"  max rel err {worst[form]:.3e}")
    assert worst["ozaki"] <= TOL
