"""GPU parity of the batched (grouped persistent) qgemm / qgemv.

REFERENCE form: every problem of the batch equals the oracle's restatement of the reference
routine on that problem bit for bit (level3.hpp:215-336 / level2.hpp:15-99 per problem), for every
tile configuration, both layouts, padded leading dimensions and strides, problems that take the
simple path (m, n, k <= 64) and the kc = 126 blocked path, and empty problems.
DD form: every entry within |got - exact| <= 2^-100 (|alpha| sum|a b| + |beta c|) + 1 ulp.
EXACT gemv: every problem equals the oracle's exact accumulation rounded once, bit for bit.
The launch sequence is captured in a graph and replayed against the eager result, and the
autotuner's choice is a valid configuration.
"""
import numpy as np
import pytest

from helpers import from_dev, pair, to_fraction

pytestmark = pytest.mark.gpu

DD_TOL = 2.0 ** -100


def _say(msg):
    print(msg, flush=True)


class Pool:
    """One device buffer holding every operand of a batch (binary128, 16-byte aligned)."""

    def __init__(self, torch, arrays):
        self.offsets = []
        off = 0
        for a in arrays:
            self.offsets.append(off)
            off += a.shape[0]
        host = np.concatenate(arrays) if arrays else np.zeros((1, 2), np.uint64)
        self.dev = torch.from_numpy(np.ascontiguousarray(host).view(np.int64)).to("cuda:0")
        self.base = self.dev.data_ptr()

    def ptr(self, idx):
        return self.base + 16 * self.offsets[idx]

    def get(self, idx, count):
        o = self.offsets[idx]
        return from_dev(self.dev[o:o + count])


def _extent(rows, cols, ld):
    return max(1, (rows - 1) * ld + cols) if rows > 0 and cols > 0 else 1


def make_gemm_batch(O, torch, layout, count, lo, hi, seed, empties=True):
    rng = np.random.default_rng(seed)
    Q_dims, arrays = [], []
    for p in range(count):
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, 3))
        if empties and p % 17 == 5:
            m = 0
        pad = p % 3
        if layout == 0:
            lda, ldb, ldc = k + pad, n + pad, n + pad
            ea, eb, ec = _extent(m, k, lda), _extent(k, n, ldb), _extent(m, n, ldc)
        else:
            lda, ldb, ldc = max(m, 1) + pad, k + pad, max(m, 1) + pad
            ea, eb, ec = _extent(k, m, lda), _extent(n, k, ldb), _extent(n, m, ldc)
        arrays += [O.fill(0, seed * 1000 + 3 * p, ea), O.fill(0, seed * 1000 + 3 * p + 1, eb),
                   O.fill(0, seed * 1000 + 3 * p + 2, ec)]
        ab = O.fill(0, seed * 7 + p, 2)
        Q_dims.append((m, n, k, lda, ldb, ldc, ab[0], ab[1]))
    pool = Pool(torch, arrays)
    import qblas_amd as Q
    desc = np.zeros(count, Q.GEMM_PROBLEM)
    for p, (m, n, k, lda, ldb, ldc, a, b) in enumerate(Q_dims):
        desc[p] = (m, n, k, lda, ldb, ldc, pool.ptr(3 * p), pool.ptr(3 * p + 1), pool.ptr(3 * p + 2), a, b, 0)
    ddesc = torch.from_numpy(desc.view(np.uint8)).to("cuda:0")
    return Q_dims, arrays, pool, ddesc


def test_gemm_batched_reference_every_config(O, Q, cuda):
    for layout in (0, 1):
        dims, arrays, pool, ddesc = make_gemm_batch(O, cuda, layout, 40, 1, 150, 11 + layout)
        want = []
        for p, (m, n, k, lda, ldb, ldc, a, b) in enumerate(dims):
            C = arrays[3 * p + 2].copy()
            if m and n and k:
                O.gemm(layout, m, n, k, a, arrays[3 * p], lda, arrays[3 * p + 1], ldb, b, C, ldc)
            want.append(C)
        C0 = [arrays[3 * p + 2] for p in range(len(dims))]
        for cfg in range(5):
            _say(f"gemm batched reference layout={layout} cfg={cfg} ({Q.gemm_batched_config(Q.MODE_REFERENCE, cfg)})")
            for p in range(len(dims)):      # restore C
                o = pool.offsets[3 * p + 2]
                pool.dev[o:o + C0[p].shape[0]] = cuda.from_numpy(C0[p].view(np.int64)).to("cuda:0")
            Q.gemm_batched_select(Q.MODE_REFERENCE, layout, len(dims), cfg)
            Q.dev_gemm_batched(Q.MODE_REFERENCE, layout, len(dims), ddesc.data_ptr())
            cuda.cuda.synchronize()
            for p in range(len(dims)):
                got = pool.get(3 * p + 2, C0[p].shape[0])
                assert np.array_equal(got, want[p]), (layout, cfg, p, dims[p][:3])
        Q.gemm_batched_select(Q.MODE_REFERENCE, layout, len(dims), 0)


def test_gemm_batched_dd_every_config(O, Q, cuda):
    for layout in (0, 1):
        dims, arrays, pool, ddesc = make_gemm_batch(O, cuda, layout, 16, 1, 130, 21 + layout)
        C0 = [arrays[3 * p + 2] for p in range(len(dims))]
        for cfg in range(4):
            _say(f"gemm batched dd layout={layout} cfg={cfg}")
            for p in range(len(dims)):
                o = pool.offsets[3 * p + 2]
                pool.dev[o:o + C0[p].shape[0]] = cuda.from_numpy(C0[p].view(np.int64)).to("cuda:0")
            Q.gemm_batched_select(Q.MODE_DD, layout, len(dims), cfg)
            Q.dev_gemm_batched(Q.MODE_DD, layout, len(dims), ddesc.data_ptr())
            cuda.cuda.synchronize()
            for p, (m, n, k, lda, ldb, ldc, a, b) in enumerate(dims):
                got = pool.get(3 * p + 2, C0[p].shape[0])
                if not (m and n and k):
                    assert np.array_equal(got, C0[p])
                    continue
                st = (lambda ld: (ld, 1)) if layout == 0 else (lambda ld: (1, ld))
                ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
                ii, jj = ii.ravel(), jj.ravel()
                A, B = arrays[3 * p], arrays[3 * p + 1]
                ex, _ = O.exact_gemm_samples(ii, jj, k, a, A, st(lda), B, st(ldb), b, C0[p], st(ldc))
                sc, _ = O.exact_gemm_samples(ii, jj, k, O.abs_array(np.array([a]))[0], O.abs_array(A), st(lda),
                                             O.abs_array(B), st(ldb), O.abs_array(np.array([b]))[0],
                                             O.abs_array(C0[p]), st(ldc))
                cs = st(ldc)
                for s in range(ii.shape[0]):
                    g = to_fraction(*pair(got[ii[s] * cs[0] + jj[s] * cs[1]]))
                    w = to_fraction(*pair(ex[s]))
                    ulp = abs(w) * 2 ** -112
                    assert abs(g - w) <= DD_TOL * to_fraction(*pair(sc[s])) + ulp, (layout, cfg, p, ii[s], jj[s])
        Q.gemm_batched_select(Q.MODE_DD, layout, len(dims), 0)


def make_gemv_batch(O, torch, layout, count, lo, hi, seed):
    import qblas_amd as Q
    rng = np.random.default_rng(seed)
    dims, arrays = [], []
    for p in range(count):
        m, n = (int(v) for v in rng.integers(lo, hi + 1, 2))
        if p % 13 == 4:
            n = 0
        incx, incy = (1, 1) if p % 3 else (2, 3)
        pad = p % 2
        lda = (n if layout == 0 else m) + pad
        lda = max(lda, 1)
        ea = _extent(m, n, lda) if layout == 0 else _extent(n, m, lda)
        arrays += [O.fill(0, seed * 1000 + 3 * p, ea), O.fill(0, seed * 1000 + 3 * p + 1, max(1, (n - 1) * incx + 1)),
                   O.fill(0, seed * 1000 + 3 * p + 2, max(1, (m - 1) * incy + 1))]
        ab = O.fill(0, seed * 5 + p, 2)
        dims.append((m, n, lda, incx, incy, ab[0], ab[1]))
    pool = Pool(torch, arrays)
    desc = np.zeros(count, Q.GEMV_PROBLEM)
    for p, (m, n, lda, incx, incy, a, b) in enumerate(dims):
        desc[p] = (m, n, lda, incx, incy, pool.ptr(3 * p), pool.ptr(3 * p + 1), pool.ptr(3 * p + 2), a, b)
    return dims, arrays, pool, torch.from_numpy(desc.view(np.uint8)).to("cuda:0")


def test_gemv_batched_reference(O, Q, cuda):
    for layout in (0, 1):
        _say(f"gemv batched reference layout={layout}")
        dims, arrays, pool, ddesc = make_gemv_batch(O, cuda, layout, 48, 1, 300, 31 + layout)
        Q.dev_gemv_batched(Q.MODE_REFERENCE, layout, len(dims), ddesc.data_ptr())
        cuda.cuda.synchronize()
        for p, (m, n, lda, incx, incy, a, b) in enumerate(dims):
            y = arrays[3 * p + 2].copy()
            if m and n:
                O.gemv(layout, m, n, a, arrays[3 * p], lda, arrays[3 * p + 1], incx, b, y, incy)
            assert np.array_equal(pool.get(3 * p + 2, y.shape[0]), y), (layout, p, m, n, incx)


def test_gemv_batched_dd(O, Q, cuda):
    for layout in (0, 1):
        _say(f"gemv batched dd layout={layout}")
        dims, arrays, pool, ddesc = make_gemv_batch(O, cuda, layout, 24, 1, 400, 41 + layout)
        Q.dev_gemv_batched(Q.MODE_DD, layout, len(dims), ddesc.data_ptr())
        cuda.cuda.synchronize()
        for p, (m, n, lda, incx, incy, a, b) in enumerate(dims):
            y0 = arrays[3 * p + 2]
            got = pool.get(3 * p + 2, y0.shape[0])
            if not (m and n):
                assert np.array_equal(got, y0)
                continue
            want = y0.copy()
            O.exact_gemv(layout, m, n, a, arrays[3 * p], lda, arrays[3 * p + 1], incx, b, want, incy)
            scale = O.abs_array(y0)
            O.exact_gemv(layout, m, n, O.abs_array(np.array([a]))[0], O.abs_array(arrays[3 * p]), lda,
                         O.abs_array(arrays[3 * p + 1]), incx, O.abs_array(np.array([b]))[0], scale, incy)
            for i in range(m):
                g, w = to_fraction(*pair(got[i * incy])), to_fraction(*pair(want[i * incy]))
                assert abs(g - w) <= DD_TOL * to_fraction(*pair(scale[i * incy])) + abs(w) * 2 ** -112, (layout, p, i)


def test_gemm_batched_graph_replays_and_tuner(O, Q, cuda):
    torch = cuda
    layout = 0
    dims, arrays, pool, ddesc = make_gemm_batch(O, torch, layout, 32, 1, 100, 51, empties=False)
    C0 = pool.dev.clone()
    # eager result
    Q.dev_gemm_batched(Q.MODE_REFERENCE, layout, len(dims), ddesc.data_ptr(),
                       torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    eager = pool.dev.clone()
    # autotuner: a valid configuration, identical results with it (REFERENCE order per entry)
    pool.dev.copy_(C0)
    best = Q.dev_gemm_batched_tune(Q.MODE_REFERENCE, layout, len(dims), ddesc.data_ptr())
    assert 0 <= best < 5 and Q.gemm_batched_config(Q.MODE_REFERENCE, best)
    _say(f"tuned config {best}: {Q.gemm_batched_config(Q.MODE_REFERENCE, best)}")
    Q.dev_gemm_batched(Q.MODE_REFERENCE, layout, len(dims), ddesc.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(pool.dev, eager)
    # graph: capture on a side stream, replay three times with eager work in between
    pool.dev.copy_(C0)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            Q.dev_gemm_batched(Q.MODE_REFERENCE, layout, len(dims), ddesc.data_ptr(), s.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        pool.dev.copy_(C0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(pool.dev, eager), rep
        _ = (pool.dev.float().sum() if rep == 1 else None)   # unrelated eager work between replays
    Q.gemm_batched_select(Q.MODE_REFERENCE, layout, len(dims), 0)


def test_gemv_batched_exact(O, Q, cuda):
    # EXACT form: every problem's y equals the oracle's exact accumulation rounded once, bit for bit
    for layout in (0, 1):
        _say(f"gemv batched exact layout={layout}")
        dims, arrays, pool, ddesc = make_gemv_batch(O, cuda, layout, 40, 1, 500, 61 + layout)
        Q.dev_gemv_batched(Q.MODE_EXACT, layout, len(dims), ddesc.data_ptr())
        cuda.cuda.synchronize()
        assert Q.dev_exact_status() == 0
        for p, (m, n, lda, incx, incy, a, b) in enumerate(dims):
            y = arrays[3 * p + 2].copy()
            if m and n:
                O.exact_gemv(layout, m, n, a, arrays[3 * p], lda, arrays[3 * p + 1], incx, b, y, incy)
            assert np.array_equal(pool.get(3 * p + 2, y.shape[0]), y), (layout, p, m, n, incx)
