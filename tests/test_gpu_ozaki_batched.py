"""GPU parity of the batched OZAKI qgemm (slice method, one slice count for the whole batch).

Each problem of the batch must equal, bit for bit, the single-problem OZAKI form run with the
same slice count (method SLICES). The batched kernels run the same per-entry computation:
  - the same scaling, digits and pass sums;
  - the same certification and exact repair.
Sampled entries of every problem must also be within 2^-104 relative error of the oracle's exact
value (exact zeros exactly). Coverage:
  - both layouts, padded leading dimensions, empty problems;
  - every slice-count configuration;
  - a problem with k > 4096 (two int32 accumulation passes);
  - rows that cancel to exact zeros and to tiny values (the exact repair path), and infinities;
  - the autotuner, which must leave the caller's C untouched and pick a valid configuration.
"""
import numpy as np
import pytest

from helpers import from_dev
from test_gpu_batched import Pool, make_gemm_batch

pytestmark = pytest.mark.gpu

TOL = 2.0 ** -104


def _say(msg):
    print(msg, flush=True)


def _strides(layout, ld):
    return (ld, 1) if layout == 0 else (1, ld)


def _restore(pool, cuda, C0, idx_of):
    for p, c in enumerate(C0):
        o = pool.offsets[idx_of(p)]
        pool.dev[o:o + c.shape[0]] = cuda.from_numpy(c.view(np.int64)).to("cuda:0")


def _check_oracle(O, layout, dims, arrays, C0, got_of, rng, per_problem=300):
    worst = 0.0
    for p, (m, n, k, lda, ldb, ldc, a, b) in enumerate(dims):
        got = got_of(p)
        if not (m and n and k):
            assert np.array_equal(got, C0[p]), ("empty problem changed", p)
            continue
        ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        ii, jj = ii.ravel(), jj.ravel()
        if ii.shape[0] > per_problem:
            sel = rng.choice(ii.shape[0], per_problem, replace=False)
            ii, jj = ii[sel], jj[sel]
        cs = _strides(layout, ldc)
        cand = got[ii * cs[0] + jj * cs[1]]
        _, rel = O.exact_gemm_samples(ii, jj, k, a, arrays[3 * p], _strides(layout, lda), arrays[3 * p + 1],
                                      _strides(layout, ldb), b, C0[p], cs, cand=cand)
        assert np.all(rel <= TOL), (p, (m, n, k), float(rel.max()))
        worst = max(worst, float(rel.max()))
    return worst


def _single_problem(Q, cuda, layout, dims, arrays, S):
    """C of every problem from the single-problem OZAKI form with S slices (method SLICES)."""
    Q.set_ozaki_method(Q.OZAKI_SLICES)
    Q.set_ozaki_slices(S)
    out = []
    try:
        for p, (m, n, k, lda, ldb, ldc, a, b) in enumerate(dims):
            dA = cuda.from_numpy(arrays[3 * p].view(np.int64)).to("cuda:0")
            dB = cuda.from_numpy(arrays[3 * p + 1].view(np.int64)).to("cuda:0")
            dC = cuda.from_numpy(arrays[3 * p + 2].view(np.int64)).to("cuda:0")
            Q.dev_gemm(Q.MODE_OZAKI, layout, m, n, k, a, dA.data_ptr(), lda, dB.data_ptr(), ldb, b,
                       dC.data_ptr(), ldc)
            cuda.cuda.synchronize()
            out.append(from_dev(dC))
    finally:
        Q.set_ozaki_slices(0)
        Q.set_ozaki_method(Q.OZAKI_AUTO)
    return out


def test_ozaki_batched_every_slice_count(O, Q, cuda):
    cfgs = [i for i in range(16) if Q.gemm_batched_config(Q.MODE_OZAKI, i)]
    assert len(cfgs) >= 4 and Q.gemm_batched_config(Q.MODE_OZAKI, len(cfgs)) is None
    rng = np.random.default_rng(5)
    for layout in (0, 1):
        dims, arrays, pool, ddesc = make_gemm_batch(O, cuda, layout, 24, 1, 150, 61 + layout)
        C0 = [arrays[3 * p + 2] for p in range(len(dims))]
        for cfg in cfgs:
            name = Q.gemm_batched_config(Q.MODE_OZAKI, cfg)
            S = int(name.split()[0])
            _say(f"ozaki batched layout={layout} cfg={cfg} ({name})")
            _restore(pool, cuda, C0, lambda p: 3 * p + 2)
            Q.gemm_batched_select(Q.MODE_OZAKI, layout, len(dims), cfg)
            Q.dev_gemm_batched(Q.MODE_OZAKI, layout, len(dims), ddesc.data_ptr())
            cuda.cuda.synchronize()
            assert Q.dev_exact_status() == 0
            assert Q.ozaki_last_slices() == S
            got = [pool.get(3 * p + 2, C0[p].shape[0]) for p in range(len(dims))]
            worst = _check_oracle(O, layout, dims, arrays, C0, lambda p: got[p], rng)
            _say(f"  max rel err {worst:.3e}, repaired entries {Q.dev_ozaki_repairs()}")
            if cfg in (0, cfgs[-1]):
                single = _single_problem(Q, cuda, layout, dims, arrays, S)
                for p in range(len(dims)):
                    assert np.array_equal(got[p], single[p]), ("batched != single-problem", layout, cfg, p, dims[p][:3])
        Q.gemm_batched_select(Q.MODE_OZAKI, layout, len(dims), 0)


def _cancel_batch(O, cuda, layout):
    """Problems whose rows cancel, with beta = 0 so that alpha AB alone decides the certification:
    A = [X, -X] along k against B = [Y; Y] makes AB exactly zero, and A = [X, -X, D] against
    B = [Y; Y; Z] with tiny D (|D| < 2^-80) leaves AB = D Z, far below the scale of the rows; plus
    a problem with k = 4200 (two accumulation passes) and beta != 0."""
    import qblas_amd as Q
    specs = [(40, 33, 64, "zero"), (50, 70, 96, "tiny"), (17, 9, 4200, "plain"), (64, 64, 130, "tiny"),
             (20, 30, 50, "inf")]
    dims, arrays = [], []
    zero = np.zeros(2, np.uint64)
    for p, (m, n, k, kind) in enumerate(specs):
        d = 8 if kind == "tiny" else 0
        h = (k - d) // 2
        X = O.fill(0, 900 + p, m * h).reshape(m, h, 2)
        Y = O.fill(0, 950 + p, h * n).reshape(h, n, 2)
        if kind in ("plain", "inf"):
            A = O.fill(0, 990 + p, m * k).reshape(m, k, 2)
            Bm = O.fill(0, 995 + p, k * n).reshape(k, n, 2)
            if kind == "inf":  # +Inf at A[3, 7] and -Inf at A[11, 0]: rows 3 and 11 become +-Inf or NaN
                A[3, 7] = (0, 0x7fff000000000000)
                A[11, 0] = (0, 0xffff000000000000)
        else:
            negX = X.copy()
            negX[..., 1] ^= np.uint64(1 << 63)
            A, Bm = [X, negX], [Y, Y]
            if d:
                A.append(O.fill(1, 980 + p, m * d, -90, -80).reshape(m, d, 2))
                Bm.append(O.fill(0, 985 + p, d * n).reshape(d, n, 2))
            A = np.concatenate(A, axis=1)
            Bm = np.concatenate(Bm, axis=0)
        C = O.fill(0, 970 + p, m * n).reshape(m, n, 2)
        if layout == 0:
            a_arr, lda = np.ascontiguousarray(A).reshape(-1, 2), k
            b_arr, ldb = np.ascontiguousarray(Bm).reshape(-1, 2), n
            c_arr, ldc = np.ascontiguousarray(C).reshape(-1, 2), n
        else:
            a_arr, lda = np.ascontiguousarray(A.transpose(1, 0, 2)).reshape(-1, 2), m
            b_arr, ldb = np.ascontiguousarray(Bm.transpose(1, 0, 2)).reshape(-1, 2), k
            c_arr, ldc = np.ascontiguousarray(C.transpose(1, 0, 2)).reshape(-1, 2), m
        arrays += [a_arr, b_arr, c_arr]
        ab = O.fill(0, 40 + p, 2)
        dims.append((m, n, k, lda, ldb, ldc, ab[0], ab[1] if kind in ("plain", "inf") else zero))
    pool = Pool(cuda, arrays)
    desc = np.zeros(len(dims), Q.GEMM_PROBLEM)
    for p, (m, n, k, lda, ldb, ldc, a, b) in enumerate(dims):
        desc[p] = (m, n, k, lda, ldb, ldc, pool.ptr(3 * p), pool.ptr(3 * p + 1), pool.ptr(3 * p + 2), a, b, 0)
    ddesc = cuda.from_numpy(desc.view(np.uint8)).to("cuda:0")
    return dims, arrays, pool, ddesc


def test_ozaki_batched_cancellation_and_long_k(O, Q, cuda):
    rng = np.random.default_rng(6)
    for layout in (0, 1):
        dims, arrays, pool, ddesc = _cancel_batch(O, cuda, layout)
        C0 = [arrays[3 * p + 2] for p in range(len(dims))]
        _say(f"ozaki batched cancellation layout={layout}")
        Q.dev_gemm_batched(Q.MODE_OZAKI, layout, len(dims), ddesc.data_ptr())
        cuda.cuda.synchronize()
        assert Q.dev_exact_status() == 0
        repaired = Q.dev_ozaki_repairs()
        assert repaired >= 40 * 33, repaired      # every entry of the exactly cancelling problem
        got = [pool.get(3 * p + 2, C0[p].shape[0]) for p in range(len(dims))]
        finite = len(dims) - 1        # the last problem holds the infinities (the oracle takes finite inputs)
        worst = _check_oracle(O, layout, dims[:finite], arrays, C0, lambda p: got[p], rng, per_problem=400)
        _say(f"  max rel err {worst:.3e}, repaired entries {repaired}")
        m, n = dims[finite][:2]
        cs = _strides(layout, dims[finite][5])
        bexp = lambda i, j: (int(got[finite][i * cs[0] + j * cs[1]][1]) >> 48) & 0x7FFF
        assert all(bexp(i, j) == 0x7FFF for i in (3, 11) for j in range(n)), "Inf rows are not Inf/NaN"
        assert all(bexp(i, j) != 0x7FFF for i in (0, 5, m - 1) for j in range(n)), "finite rows hit"
        S = Q.ozaki_last_slices()
        single = _single_problem(Q, cuda, layout, dims, arrays, S)
        for p in range(len(dims)):
            assert np.array_equal(got[p], single[p]), ("batched != single-problem", layout, p, dims[p][:3])


def test_ozaki_batched_tuner(O, Q, cuda):
    layout = 0
    dims, arrays, pool, ddesc = make_gemm_batch(O, cuda, layout, 48, 8, 200, 71, empties=False)
    before = pool.dev.clone()
    best = Q.dev_gemm_batched_tune(Q.MODE_OZAKI, layout, len(dims), ddesc.data_ptr())
    cuda.cuda.synchronize()
    assert Q.gemm_batched_config(Q.MODE_OZAKI, best), best
    assert cuda.equal(pool.dev, before), "the tuner wrote the caller's matrices"
    _say(f"tuned slice count: {Q.gemm_batched_config(Q.MODE_OZAKI, best)}")
    Q.dev_gemm_batched(Q.MODE_OZAKI, layout, len(dims), ddesc.data_ptr())
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() == 0
    C0 = [arrays[3 * p + 2] for p in range(len(dims))]
    got = [pool.get(3 * p + 2, C0[p].shape[0]) for p in range(len(dims))]
    _check_oracle(O, layout, dims, arrays, C0, lambda p: got[p], np.random.default_rng(7), per_problem=200)
    Q.gemm_batched_select(Q.MODE_OZAKI, layout, len(dims), 0)


def test_ozaki_batched_graph_replay(O, Q, cuda):
    """An eager call, then the same call captured in a graph: every replay (with eager work
    between replays) gives the eager result bit for bit; a replay after a descriptor's shape
    changed computes nothing and reports it; restoring the shape makes replays correct again."""
    layout = 0
    dims, arrays, pool, ddesc = make_gemm_batch(O, cuda, layout, 40, 8, 160, 83, empties=True)
    count = len(dims)
    C0 = [arrays[3 * p + 2] for p in range(count)]
    desc_host = ddesc.clone()
    _say("batched OZAKI: eager call")
    Q.dev_gemm_batched(Q.MODE_OZAKI, layout, count, ddesc.data_ptr())
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() == 0
    want = [pool.get(3 * p + 2, C0[p].shape[0]) for p in range(count)]
    _check_oracle(O, layout, dims, arrays, C0, lambda p: want[p], np.random.default_rng(8), per_problem=60)
    _restore(pool, cuda, C0, lambda p: 3 * p + 2)
    cuda.cuda.synchronize()
    _say("batched OZAKI: capture")
    g = cuda.cuda.CUDAGraph()
    s = cuda.cuda.Stream()
    with cuda.cuda.stream(s):
        with cuda.cuda.graph(g, stream=s):
            Q.dev_gemm_batched(Q.MODE_OZAKI, layout, count, ddesc.data_ptr(), s.cuda_stream)
    cuda.cuda.synchronize()
    for r in range(3):
        _say(f"batched OZAKI: replay {r}")
        _restore(pool, cuda, C0, lambda p: 3 * p + 2)
        cuda.cuda.synchronize()
        g.replay()
        cuda.cuda.synchronize()
        assert Q.dev_exact_status() == 0
        for p in range(count):
            assert np.array_equal(pool.get(3 * p + 2, C0[p].shape[0]), want[p]), (r, p)
        scratch = cuda.ones(1 << 20, device="cuda:0") * 3.0   # eager work between replays
        assert float(scratch.sum()) == 3.0 * (1 << 20)
    _say("batched OZAKI: replay after a shape change")
    desc = desc_host.cpu().numpy().view(Q.GEMM_PROBLEM).copy()
    q = next(p for p, d in enumerate(dims) if d[0] > 1 and d[1] and d[2])
    desc[q]["m"] -= 1
    ddesc.copy_(cuda.from_numpy(desc.view(np.uint8)).to("cuda:0"))
    _restore(pool, cuda, C0, lambda p: 3 * p + 2)
    cuda.cuda.synchronize()
    g.replay()
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() != 0
    for p in range(count):
        assert np.array_equal(pool.get(3 * p + 2, C0[p].shape[0]), C0[p]), ("C written by a stale replay", p)
    ddesc.copy_(desc_host)
    _restore(pool, cuda, C0, lambda p: 3 * p + 2)
    cuda.cuda.synchronize()
    g.replay()
    cuda.cuda.synchronize()
    assert Q.dev_exact_status() == 0
    for p in range(count):
        assert np.array_equal(pool.get(3 * p + 2, C0[p].shape[0]), want[p]), ("after restore", p)
    del g


def test_ozaki_batched_rejected_for_gemv(Q, cuda):
    desc = cuda.zeros(64, dtype=cuda.uint8, device="cuda:0")
    with pytest.raises(RuntimeError):
        Q.dev_gemv_batched(Q.MODE_OZAKI, 0, 1, desc.data_ptr())


def _quantiles(v):
    v = np.sort(np.asarray(v, dtype=np.float64))
    q = lambda f: float(v[min(len(v) - 1, int(f * len(v)))])
    return {"median": q(0.5), "p90": q(0.9), "p99": q(0.99), "max": float(v[-1])}


def test_batched_full_size_error_distribution(O, Q, cuda):
    """The batched workload at full size: 4096 qgemm problems (m, n, k uniform in [8, 256]) and
    16384 qgemv problems (m, n uniform in [32, 1024]), RowMajor, entries and per-problem alpha, beta
    full-mantissa uniform [-1, 1) (device generator).  Per problem, the largest relative error of
    sampled entries against the oracle's exact value; its distribution over the problems is
    printed for every form.  OZAKI gemm stays within 2^-104 and EXACT gemv is correctly rounded
    on every sample."""
    torch = cuda
    rng = np.random.default_rng(77)
    # ---- gemm batch
    count = 4096
    dims = rng.integers(8, 257, size=(count, 3))
    sizes = dims[:, 0] * dims[:, 2] + dims[:, 2] * dims[:, 1] + dims[:, 0] * dims[:, 1]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pool = torch.empty((int(offs[-1]), 2), dtype=torch.int64, device="cuda:0")
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 77, pool.shape[0], pool.data_ptr())
    ab = torch.empty((2 * count, 2), dtype=torch.int64, device="cuda:0")
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 78, ab.shape[0], ab.data_ptr())
    torch.cuda.synchronize()
    abh = from_dev(ab)
    host = from_dev(pool)
    desc = np.zeros(count, Q.GEMM_PROBLEM)
    base = pool.data_ptr()
    S_PER = 4
    samples = []
    for p in range(count):
        m, n, k = (int(v) for v in dims[p])
        o = int(offs[p])
        desc[p] = (m, n, k, k, n, n, base + 16 * o, base + 16 * (o + m * k), base + 16 * (o + m * k + k * n),
                   abh[2 * p], abh[2 * p + 1], 0)
        samples.append((rng.integers(0, m, S_PER), rng.integers(0, n, S_PER)))
    ddesc = torch.from_numpy(desc.view(np.uint8)).to("cuda:0")
    cpos = np.concatenate([offs[p] + dims[p, 0] * dims[p, 2] + dims[p, 2] * dims[p, 1] +
                           samples[p][0] * dims[p, 1] + samples[p][1] for p in range(count)])
    cidx = torch.from_numpy(cpos).to("cuda:0")
    pool0 = pool.clone()
    for form in ("ozaki", "dd", "reference"):
        mode = Q.MODES[form]
        pool.copy_(pool0)
        if form == "ozaki":
            best = Q.dev_gemm_batched_tune(mode, 0, count, ddesc.data_ptr())
            _say(f"4096 gemm problems: tuned {Q.gemm_batched_config(mode, best)}")
        Q.dev_gemm_batched(mode, 0, count, ddesc.data_ptr())
        torch.cuda.synchronize()
        if form == "ozaki":
            assert Q.dev_exact_status() == 0
        got = from_dev(pool[cidx])
        worst = []
        for p in range(count):
            m, n, k = (int(v) for v in dims[p])
            o = int(offs[p])
            A, B = host[o:o + m * k], host[o + m * k:o + m * k + k * n]
            C0 = host[o + m * k + k * n:o + m * k + k * n + m * n]
            _, rel = O.exact_gemm_samples(samples[p][0], samples[p][1], k, abh[2 * p], A, (k, 1), B, (n, 1),
                                          abh[2 * p + 1], C0, (n, 1), cand=got[S_PER * p:S_PER * (p + 1)])
            worst.append(float(rel.max()))
        qs = _quantiles(worst)
        _say(f"4096 gemm problems, {form}: per-problem max rel err over {S_PER} samples: " +
             ", ".join(f"{key} {val:.3e}" for key, val in qs.items()))
        if form == "ozaki":
            assert qs["max"] <= TOL
    del pool, pool0, host, cidx
    torch.cuda.empty_cache()

    # ---- gemv batch (A is 73 GB on the device; one sampled row per problem is copied back)
    count = 16384
    mn = rng.integers(32, 1025, size=(count, 2))
    asz = mn[:, 0] * mn[:, 1]
    offs = np.concatenate([[0], np.cumsum(asz + mn[:, 1] + mn[:, 0])]).astype(np.int64)
    pool = torch.empty((int(offs[-1]), 2), dtype=torch.int64, device="cuda:0")
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 79, pool.shape[0], pool.data_ptr())
    ab = torch.empty((2 * count, 2), dtype=torch.int64, device="cuda:0")
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 80, ab.shape[0], ab.data_ptr())
    torch.cuda.synchronize()
    abh = from_dev(ab)
    desc = np.zeros(count, Q.GEMV_PROBLEM)
    base = pool.data_ptr()
    rows, arow, xv = rng.integers(0, mn[:, 0]), [], []
    for p in range(count):
        m, n = int(mn[p, 0]), int(mn[p, 1])
        o = int(offs[p])
        desc[p] = (m, n, n, 1, 1, base + 16 * o, base + 16 * (o + m * n), base + 16 * (o + m * n + n),
                   abh[2 * p], abh[2 * p + 1])
        i = int(rows[p])
        arow.append(from_dev(pool[o + i * n:o + (i + 1) * n]))
        xv.append(from_dev(pool[o + m * n:o + m * n + n]))
    ddesc = torch.from_numpy(desc.view(np.uint8)).to("cuda:0")
    ypos = np.concatenate([offs[p] + asz[p] + mn[p, 1] + np.arange(mn[p, 0]) for p in range(count)])
    yidx = torch.from_numpy(ypos).to("cuda:0")
    y0 = pool[yidx].clone()
    spos = torch.from_numpy((offs[:-1] + asz + mn[:, 1] + rows).astype(np.int64)).to("cuda:0")
    y0s = from_dev(pool[spos])
    zero_i = np.zeros(1, np.int64)
    for form in ("exact", "dd", "reference"):
        mode = Q.MODES[form]
        pool[yidx] = y0
        Q.dev_gemv_batched(mode, 0, count, ddesc.data_ptr())
        torch.cuda.synchronize()
        if form == "exact":
            assert Q.dev_exact_status() == 0
        got = from_dev(pool[spos])
        worst, mism = [], 0
        for p in range(count):
            n = int(mn[p, 1])
            rounded, rel = O.exact_gemm_samples(zero_i, zero_i, n, abh[2 * p], arow[p], (n, 1), xv[p], (1, 0),
                                                abh[2 * p + 1], y0s[p:p + 1], (0, 0), cand=got[p:p + 1])
            worst.append(float(rel[0]))
            mism += int(not np.array_equal(rounded[0], got[p]))
        qs = _quantiles(worst)
        _say(f"16384 gemv problems, {form}: rel err of one sampled row per problem: " +
             ", ".join(f"{key} {val:.3e}" for key, val in qs.items()) + f"; not correctly rounded: {mism}")
        if form == "exact":
            assert mism == 0
