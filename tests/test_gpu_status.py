"""Per-call status words and per-stream workspaces.

- The status of an EXACT- or OZAKI-form device call stays attached to that call: a later
  REFERENCE call on the same stream (which reuses the stream's workspace) must neither make a
  host-pointer call report an error nor be reported as that call's status.
- Batched calls on two streams at once, of different counts, each with its own workspace: every
  result is bit-identical to the same call made alone.
"""
import ctypes

import numpy as np
import pytest

from helpers import dev_empty, from_dev, pair, to_dev

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_exact_status_then_reference_then_host_dot(O, Q, cuda):
    n = 5000
    x, y = O.fill(0, 31, n), O.fill(0, 32, n)
    dx, dy, dr = to_dev(x), to_dev(y), dev_empty(1)
    s = cuda.cuda.Stream()
    _say("exact dot on a user stream, then a reference dot on the same stream, then quadblas_qdot")
    Q.dev_dot(Q.MODE_EXACT, n, dx.data_ptr(), 1, dy.data_ptr(), 1, 1, dr.data_ptr(), s.cuda_stream)
    Q.dev_dot(Q.MODE_REFERENCE, n, dx.data_ptr(), 1, dy.data_ptr(), 1, 8, dr.data_ptr(), s.cuda_stream)
    s.synchronize()
    assert pair(from_dev(dr)[0]) == O.to_pair(O.dot(n, x, 1, y, 1, 8))
    T = Q.get_num_threads()
    got = Q.lib().quadblas_qdot(n, x.ctypes.data_as(ctypes.c_void_p), 1, y.ctypes.data_as(ctypes.c_void_p), 1)
    assert Q.last_error() == ""
    assert got == O.to_double(O.dot(n, x, 1, y, 1, T))
    # the EXACT call's workspace was reused by the REFERENCE call: its status is no longer known
    Q.lib().qblas_clear_error()
    Q.dev_dot(Q.MODE_EXACT, n, dx.data_ptr(), 1, dy.data_ptr(), 1, 1, dr.data_ptr(), s.cuda_stream)
    Q.dev_dot(Q.MODE_REFERENCE, n, dx.data_ptr(), 1, dy.data_ptr(), 1, 8, dr.data_ptr(), s.cuda_stream)
    assert Q.dev_exact_status(s.cuda_stream) == -1
    Q.lib().qblas_clear_error()
    # with no call in between it is reported (0: exact)
    Q.dev_dot(Q.MODE_EXACT, n, dx.data_ptr(), 1, dy.data_ptr(), 1, 1, dr.data_ptr(), s.cuda_stream)
    assert Q.dev_exact_status(s.cuda_stream) == 0
    assert Q.dev_exact_wide_items(s.cuda_stream) == 0
    assert pair(from_dev(dr)[0]) == O.exact_dot(x, y)[0]


def test_ozaki_repairs_not_stale(O, Q, cuda):
    g = 64
    A, B, C = O.fill(1, 41, g * g, -20, 20), O.fill(1, 42, g * g, -20, 20), O.fill(0, 43, g * g)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    s = cuda.cuda.Stream()
    Q.dev_gemm(Q.MODE_OZAKI, 0, g, g, g, 1.0, dA.data_ptr(), g, dB.data_ptr(), g, 0.0, dC.data_ptr(), g,
               s.cuda_stream)
    assert Q.dev_ozaki_repairs(s.cuda_stream) >= 0
    # a DD gemm on the same stream reuses the workspace: the repair count is no longer known
    Q.dev_gemm(Q.MODE_DD, 0, g, g, g, 1.0, dA.data_ptr(), g, dB.data_ptr(), g, 0.0, dC.data_ptr(), g, s.cuda_stream)
    assert Q.dev_ozaki_repairs(s.cuda_stream) == -1


def test_batched_two_streams_each_equal_to_alone(O, Q, cuda):
    from test_gpu_batched import make_gemm_batch, make_gemv_batch
    layout = 0
    g_dims, g_arrays, g_pool, g_desc = make_gemm_batch(O, cuda, layout, 57, 1, 90, 71, empties=False)
    v_dims, v_arrays, v_pool, v_desc = make_gemv_batch(O, cuda, layout, 131, 1, 300, 72)

    def reset():
        for idx in range(len(g_dims)):
            o = g_pool.offsets[3 * idx + 2]
            c = g_arrays[3 * idx + 2]
            g_pool.dev[o:o + c.shape[0]] = cuda.from_numpy(c.view(np.int64)).to("cuda:0")
        for idx in range(len(v_dims)):
            o = v_pool.offsets[3 * idx + 2]
            c = v_arrays[3 * idx + 2]
            v_pool.dev[o:o + c.shape[0]] = cuda.from_numpy(c.view(np.int64)).to("cuda:0")
        cuda.cuda.synchronize()

    def outputs():
        return ([g_pool.get(3 * i + 2, g_arrays[3 * i + 2].shape[0]) for i in range(len(g_dims))],
                [v_pool.get(3 * i + 2, v_arrays[3 * i + 2].shape[0]) for i in range(len(v_dims))])

    for mode in (Q.MODE_REFERENCE, Q.MODE_DD):
        _say(f"batched gemm and gemv alone, mode={mode}")
        reset()
        Q.dev_gemm_batched(mode, layout, len(g_dims), g_desc.data_ptr())
        cuda.cuda.synchronize()
        Q.dev_gemv_batched(mode, layout, len(v_dims), v_desc.data_ptr())
        cuda.cuda.synchronize()
        alone_g, alone_v = outputs()
        sa, sb = cuda.cuda.Stream(), cuda.cuda.Stream()
        for rnd in range(3):
            _say(f"batched gemm and gemv on two streams at once, mode={mode}, round {rnd}")
            reset()
            Q.dev_gemm_batched(mode, layout, len(g_dims), g_desc.data_ptr(), sa.cuda_stream)
            Q.dev_gemv_batched(mode, layout, len(v_dims), v_desc.data_ptr(), sb.cuda_stream)
            cuda.cuda.synchronize()
            got_g, got_v = outputs()
            for a, b in zip(got_g, alone_g):
                assert np.array_equal(a, b)
            for a, b in zip(got_v, alone_v):
                assert np.array_equal(a, b)
            sa, sb = sb, sa


def test_dot_q_writes_nan_on_failure(O, Q, cuda):
    # a negative stride is rejected by the C ABI: the binary128 result must be a NaN, not a stale
    # or zero value
    res = np.array([123, 456], dtype=np.uint64)
    x = O.fill(0, 1, 10)
    Q.lib().quadblas_dot_q(10, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(-1 & ((1 << 64) - 1)),
                           x.ctypes.data_as(ctypes.c_void_p), 1, res.ctypes.data_as(ctypes.c_void_p))
    err = Q.last_error()
    Q.lib().qblas_clear_error()
    assert err != ""
    hi = int(res[1])
    assert ((hi >> 48) & 0x7FFF) == 0x7FFF and (((hi & ((1 << 48) - 1)) | int(res[0])) != 0)
