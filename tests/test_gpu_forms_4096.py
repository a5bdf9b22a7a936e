"""The qgemm forms at the headline size, 4096^3 RowMajor, alpha = 1, beta = 0, entries with random
sign, log-uniform magnitude over [2^-30, 2^30) and a full mantissa, on 4096 fixed sampled entries:

  plain (REFERENCE order): bit for bit against the oracle's per-entry restatement of the
      reference's blocked gemm (level3.hpp:215-336, kc = 126; a 1 x 1 x 4096 problem takes the same
      per-entry path: fresh chains over each k block, C = fma(alpha, t, beta_eff C));
  DD: within |got - exact| <= 2^-100 (|alpha| sum |a b| + |beta c|) + 1 ulp of the exact value;
  OZAKI (slice method; the CRT method is tests/test_gpu_ozaki_crt.py::test_crt_4096_cube_sampled):
      <= 2^-104 relative error against the exact value.
"""
import numpy as np
import pytest

from helpers import from_dev, to_fraction

pytestmark = pytest.mark.gpu

N = 4096
DD_TOL = 2.0 ** -100


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def data(O):
    A = O.fill(1, 11, N * N, -30, 30)
    B = O.fill(1, 12, N * N, -30, 30)
    rng = np.random.default_rng(2024)
    ii, jj = rng.integers(0, N, 4096), rng.integers(0, N, 4096)
    return A, B, ii, jj


def _run(Q, cuda, mode, A, B):
    dA = cuda.from_numpy(A.view(np.int64)).to("cuda:0")
    dB = cuda.from_numpy(B.view(np.int64)).to("cuda:0")
    dC = cuda.zeros((N * N, 2), dtype=cuda.int64, device="cuda:0")
    Q.dev_gemm(mode, 0, N, N, N, 1.0, dA.data_ptr(), N, dB.data_ptr(), N, 0.0, dC.data_ptr(), N)
    cuda.cuda.synchronize()
    out = from_dev(dC)
    del dA, dB, dC
    cuda.cuda.empty_cache()
    return out


def test_plain_4096_bit_exact_on_samples(O, Q, cuda, data):
    A, B, ii, jj = data
    _say("plain qgemm 4096^3 (REFERENCE order)")
    C = _run(Q, cuda, Q.MODE_REFERENCE, A, B)
    c = np.zeros((1, 2), dtype=np.uint64)
    bad = []
    for s, (i, j) in enumerate(zip(ii, jj)):
        c[:] = 0
        O.gemm(0, 1, 1, N, 1.0, np.ascontiguousarray(A[i * N:(i + 1) * N]), N, np.ascontiguousarray(B[j::N]), 1,
               0.0, c, 1)
        if not np.array_equal(c[0], C[i * N + j]):
            bad.append((int(i), int(j)))
    assert not bad, bad[:10]


def test_dd_4096_within_bound_on_samples(O, Q, cuda, data):
    A, B, ii, jj = data
    _say("dd qgemm 4096^3")
    C = _run(Q, cuda, Q.MODE_DD, A, B)
    cand = np.ascontiguousarray(C[ii * N + jj])
    exact, rel = O.exact_gemm_samples(ii, jj, N, 1.0, A, (N, 1), B, (N, 1), 0.0, None, (N, 1), cand=cand)
    scale, _ = O.exact_gemm_samples(ii, jj, N, 1.0, O.abs_array(A), (N, 1), O.abs_array(B), (N, 1), 0.0, None, (N, 1))
    bad = 0
    for g, w, sc in zip(cand, exact, scale):
        gv, wv, sv = to_fraction(*g), to_fraction(*w), to_fraction(*sc)
        ulp = abs(wv) * (2 ** -112)
        bad += int(abs(gv - wv) > DD_TOL * sv + ulp)
    _say(f"  dd max rel {rel.max():.3e}, median {np.median(rel):.3e}")
    assert bad == 0


def test_ozaki_slices_4096_on_samples(O, Q, cuda, data):
    A, B, ii, jj = data
    _say("ozaki slice method qgemm 4096^3")
    prev = Q.get_ozaki_method()
    Q.set_ozaki_method(Q.OZAKI_SLICES)
    try:
        C = _run(Q, cuda, Q.MODE_OZAKI, A, B)
        assert Q.dev_exact_status() == 0
    finally:
        Q.set_ozaki_method(prev)
    cand = np.ascontiguousarray(C[ii * N + jj])
    _, rel = O.exact_gemm_samples(ii, jj, N, 1.0, A, (N, 1), B, (N, 1), 0.0, None, (N, 1), cand=cand)
    _say(f"  slices max rel {rel.max():.3e}")
    assert rel.max() <= 2.0 ** -104
