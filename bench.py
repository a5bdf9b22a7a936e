"""Benchmark of the QuadBLAS binary128 hot path on one MI355X.

Headline (`value`): qgemm M = N = K = 4096, RowMajor, alpha = 1, beta = 0, entries with random sign,
magnitude log-uniform over [2^-30, 2^30) and a full 113-bit mantissa (seeded device generator),
in the fastest form that meets 2^-104 relative error: OZAKI (operands scaled per row / column and
rounded to integers, integer products on the int8 matrix cores with exact int32 accumulation,
exact recombination; every entry is certified by an a-priori bound to <= 2^-105.9 relative error
or recomputed exactly), with the library's automatic method: for this workload the CRT method
(residues modulo ~40 moduli <= 256, one int8 GEMM per modulus) rather than the slice method
(S(S+1)/2 slice-pair products; reported as qgemm_forms.ozaki_slices).  value = 2 N^3 / (ms per call), quad-equivalent
GFLOP/s (the reference's own GFLOPS formula, benchmarks/benchmark.cpp:198-202).  Inputs are resident
in HBM before the timed region; one step = one qgemm call.

Also reported:
  qgemm_forms     the same problem in the plain soft-binary128 form (the reference's summation
                  order and blocking, one correctly rounded FMA at a time), the DD form (double-
                  double on FP64 FMA; fp64: its FP64 rate from the kernel's 30 flops / 27 FP64
                  instructions per multiply-add, against 78.6 TFLOP/s) and the OZAKI form.
  memory_bound    qdot n = 2^24 and qgemv 8192 x 8192 RowMajor / ColMajor (alpha = 1.5,
                  beta = -0.5, entries a + b 2^-60) in the DD and EXACT forms: GB/s against the
                  8 TB/s HBM peak.  Algorithmic bytes: dot 32 n, gemv 16 (m n + n + 2 m).
  roofline        the product kernel of the headline call: int8 ops 2 M N K P per launch (P = the
                  number of moduli for the CRT method, S(S+1)/2 slice pairs for the slice method)
                  over its launch duration, timed with HIP events recorded by the library on the
                  call's stream, against 5000 TOP/s dense int8.
                  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE runs
                  (profiles/ozaki_traffic.json) when present.
  batched         4096 qgemm problems (m, n, k uniform in [8, 256]) and 16384 qgemv problems
                  (m, n uniform in [32, 1024]) with per-problem alpha, beta, each batch one grouped
                  persistent launch (tile configuration from the autotuner) replayed from a
                  captured graph: problems/s, quad GFLOP/s or GB/s, REFERENCE and DD forms; the
                  qgemm batch also in the OZAKI form (slice count from the autotuner, eager calls:
                  the form synchronizes once per call to size its workspace).
  host_abi        the reference's C ABI with host buffers (pageable and pinned): the same qgemm,
                  qgemv RowMajor and qdot with their PCIe copies in and out, wall-clock per call.
  ill_conditioned qgemm 8192 x 8192 x 16384 ColMajor whose entries cancel to ~1e-20 of the
                  largest products: OZAKI (the library's automatic choice), DD and plain, with the
                  call's HBM (operands + workspace) and the algorithmic HBM rate; CPU slab baseline.
  config1_qgemv_1500  the reference's own CPU benchmark case (README.md:23, 1.6 GFLOP/s): plain
                  qgemv RowMajor 1500 x 1500 in the REFERENCE form on the GPU (device-resident and
                  through the C ABI with host buffers), the oracle's restatement on the host cores
                  (oracle FMA and libquadmath fmaq), and a bit-for-bit check of both GPU results.
  qgemm_forms[*].error  every form's max / median relative error on 4096 fixed sampled entries
                  against the exact value and against the REFERENCE form's result; the plain form's
                  samples checked bit for bit against the oracle's per-entry restatement.
  cpu_baseline    the oracle's restatement of the reference qgemm (level3.hpp blocking, soft
                  binary128 FMA, OpenMP over the host cores) on a 256-row slab of the same
                  4096^3 problem, scaled to GFLOP/s (per-element cost is uniform); with the CPU model
                  and a libquadmath-fmaq variant.  The batched legs get a serial loop over their
                  first 256 problems, extrapolated.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain] [--size N]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 vector (an FMA counts 2 flops)
# k_gemm_dd's loop body per multiply-add (the library's gfx950 code): TwoProd (1 v_mul_f64 +
# 1 v_fma_f64), the two cross terms (2 v_fma_f64) and the accurate double-double addition
# (23 v_add_f64): 27 FP64 instructions, 30 flops
DD_GEMM_FP64_INSTS_PER_MAC = 27
DD_GEMM_FP64_FLOPS_PER_MAC = 30
METRIC = "quad-equiv GFLOP/s, qgemm 4096³ at ≤2^-104 rel err; qdot/qgemv HBM GB/s"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--size", type=int, default=4096, help="qgemm M = N = K")
    p.add_argument("--slices", type=int, default=0, help="OZAKI slice count (0 = library choice)")
    p.add_argument("--method", choices=("auto", "slices", "crt"), default="auto",
                   help="OZAKI method of the headline call (auto: the library's choice)")
    p.add_argument("--crt-bits", type=int, default=0, help="OZAKI CRT bits (0 = library choice)")
    p.add_argument("--no-slices-form", action="store_true",
                   help="skip the separate slice-method line when the headline uses the CRT method")
    p.add_argument("--no-plain", action="store_true", help="skip the plain soft-binary128 qgemm")
    p.add_argument("--no-memory", action="store_true", help="skip qdot / qgemv")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    p.add_argument("--no-batched", action="store_true", help="skip the batched problems")
    p.add_argument("--no-ill", action="store_true", help="skip the ill-conditioned 8192x8192x16384 qgemm")
    p.add_argument("--no-host", action="store_true", help="skip the host-buffer C ABI calls")
    p.add_argument("--no-ill-plain", action="store_true",
                   help="skip the plain form (~10 s) of the ill-conditioned qgemm")
    p.add_argument("--batch-gemm", type=int, default=4096, help="problems in the batched qgemm")
    p.add_argument("--batch-gemv", type=int, default=16384, help="problems in the batched qgemv")
    p.add_argument("--cpu-rows", type=int, default=256, help="rows of C in the CPU qgemm slab")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first run one device copy of 1 GiB (a known-size reference for HBM counters)")
    p.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                   help="write the per-kernel HBM traffic of two rocprofv3 counter runs (FETCH_SIZE, "
                        "WRITE_SIZE; --copy-calibration --steps 1 --warmup 0) as JSON to stdout, then exit")
    p.add_argument("--n-dot", type=int, default=1 << 24)
    p.add_argument("--m", type=int, default=8192)
    return p.parse_args()


def events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


ERR_SAMPLES = 4096       # fixed sampled entries of C for the error statement of every qgemm form
ERR_SEED = 2024


def sample_index(N):
    """The fixed sampled entries (i, j) of the N^3 qgemm (seeded), and their flat RowMajor indices."""
    import numpy as np
    rng = np.random.default_rng(ERR_SEED)
    ii, jj = rng.integers(0, N, ERR_SAMPLES), rng.integers(0, N, ERR_SAMPLES)
    return ii, jj, ii * N + jj


def gather(torch, C, flat):
    """binary128 entries C[flat] of a device (N*N, 2) int64 tensor, as a host (len, 2) uint64 array."""
    import numpy as np
    idx = torch.from_numpy(np.ascontiguousarray(flat, dtype=np.int64)).to(C.device)
    return C.index_select(0, idx).cpu().numpy().view(np.uint64).copy()


def bench_ozaki(Q, torch, dev, stream, N, steps, warmup, slices, method=0, crt_bits=0):
    sh = stream.cuda_stream
    A = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    B = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    C = torch.zeros((N * N, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 11, N * N, A.data_ptr(), -30, 30, stream=sh)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 12, N * N, B.data_ptr(), -30, 30, stream=sh)
    Q.set_ozaki_slices(slices)
    Q.set_ozaki_method(method)
    Q.set_ozaki_crt_bits(crt_bits)

    def call():
        Q.dev_gemm(Q.MODE_OZAKI, 0, N, N, N, 1.0, A.data_ptr(), N, B.data_ptr(), N, 0.0, C.data_ptr(), N, sh)

    Q.set_phase_timing(True)
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    step_ms, phases = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s_ev, e_ev = events(torch)
        s_ev.record(stream)
        call()
        e_ev.record(stream)
        phases.append(Q.phase_times())       # synchronizes with the call's last phase event
        e_ev.synchronize()
        step_ms.append(s_ev.elapsed_time(e_ev))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    Q.set_phase_timing(False)
    S = Q.ozaki_last_slices()
    used = Q.ozaki_last_method()
    bits, moduli = Q.ozaki_last_crt()
    repairs = Q.dev_ozaki_repairs(sh)
    status = Q.dev_exact_status(sh)
    Q.set_ozaki_slices(0)
    Q.set_ozaki_method(Q.OZAKI_AUTO)
    Q.set_ozaki_crt_bits(0)
    ms = statistics.mean(step_ms)
    ph = {k: round(statistics.mean(p[k] for p in phases), 4) for k in phases[0]}
    samples = gather(torch, C, sample_index(N)[2])
    del A, B, C
    torch.cuda.empty_cache()
    return {"ms": ms, "wall_s": wall, "method": "crt" if used == Q.OZAKI_CRT else "slices", "slices": S,
            "crt_bits": bits, "moduli": moduli, "repairs": int(repairs), "status": int(status), "phases_ms": ph,
            "quad_GFLOPs": 2.0 * N ** 3 / (ms * 1e-3) / 1e9, "samples": samples}


def ozaki_form(oz):
    f = {"ms": round(oz["ms"], 3), "quad_GFLOPs": round(oz["quad_GFLOPs"], 1), "method": oz["method"]}
    if oz["method"] == "crt":
        f.update(crt_bits=oz["crt_bits"], moduli=oz["moduli"])
    else:
        f["slices"] = oz["slices"]
    f.update(repaired_entries=oz["repairs"], phases_ms=oz["phases_ms"], samples=oz["samples"])
    return f


def bench_form_once(Q, torch, dev, stream, N, mode):
    """One warm-up and one timed qgemm N^3 call of the given form (same data as the headline)."""
    sh = stream.cuda_stream
    A = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    B = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    C = torch.zeros((N * N, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 11, N * N, A.data_ptr(), -30, 30, stream=sh)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 12, N * N, B.data_ptr(), -30, 30, stream=sh)

    def call():
        Q.dev_gemm(mode, 0, N, N, N, 1.0, A.data_ptr(), N, B.data_ptr(), N, 0.0, C.data_ptr(), N, sh)

    call()
    torch.cuda.synchronize()
    s_ev, e_ev = events(torch)
    s_ev.record(stream)
    call()
    e_ev.record(stream)
    torch.cuda.synchronize()
    ms = s_ev.elapsed_time(e_ev)
    samples = gather(torch, C, sample_index(N)[2])
    return {"ms": round(ms, 3), "quad_GFLOPs": round(2.0 * N ** 3 / (ms * 1e-3) / 1e9, 2), "samples": samples}


def bench_plain(Q, torch, dev, stream, N):
    sh = stream.cuda_stream
    A = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    B = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    C = torch.zeros((N * N, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 11, N * N, A.data_ptr(), -30, 30, stream=sh)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 12, N * N, B.data_ptr(), -30, 30, stream=sh)
    torch.cuda.synchronize()
    s_ev, e_ev = events(torch)
    s_ev.record(stream)
    Q.dev_gemm(Q.MODE_REFERENCE, 0, N, N, N, 1.0, A.data_ptr(), N, B.data_ptr(), N, 0.0, C.data_ptr(), N, sh)
    e_ev.record(stream)
    torch.cuda.synchronize()
    ms = s_ev.elapsed_time(e_ev)
    samples = gather(torch, C, sample_index(N)[2])
    return {"ms": round(ms, 3), "quad_GFLOPs": round(2.0 * N ** 3 / (ms * 1e-3) / 1e9, 2), "samples": samples}


def bench_config1(Q, torch, dev, stream, m, steps):
    """The reference's own CPU benchmark case: plain qgemv RowMajor m x m, y := 1 A x + 0 y, entries
    a + b 2^-60 (full mantissa), y0 = 0, in the REFERENCE form (the reference's order, bit-identical
    to its result).  Device-resident calls timed back to back between events, and the same call
    through the reference's C ABI with host buffers (quadblas_qgemv: PCIe copies in and out)."""
    import numpy as np
    sh = stream.cuda_stream
    A = torch.empty((m * m, 2), dtype=torch.int64, device=dev)
    x = torch.empty((m, 2), dtype=torch.int64, device=dev)
    y = torch.zeros((m, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 42, m * m, A.data_ptr(), stream=sh)
    Q.dev_fill(Q.FILL_UNIFORM_FULL, 43, m, x.data_ptr(), stream=sh)

    def call():
        Q.dev_gemv(Q.MODE_REFERENCE, 0, m, m, 1.0, A.data_ptr(), m, x.data_ptr(), 1, 0.0, y.data_ptr(), 1, sh)

    flops = 2.0 * m * m

    def timed(mode):
        def run():
            Q.dev_gemv(mode, 0, m, m, 1.0, A.data_ptr(), m, x.data_ptr(), 1, 0.0, y.data_ptr(), 1, sh)
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        s_ev, e_ev = events(torch)
        s_ev.record(stream)
        for _ in range(steps):
            run()
        e_ev.record(stream)
        torch.cuda.synchronize()
        ms_ = s_ev.elapsed_time(e_ev) / steps
        return {"ms": round(ms_, 4), "quad_GFLOPs": round(flops / (ms_ * 1e-3) / 1e9, 2)}

    other = {"exact": timed(Q.MODE_EXACT), "dd": timed(Q.MODE_DD)}
    ms = timed(Q.MODE_REFERENCE)["ms"]
    y.zero_()
    call()
    torch.cuda.synchronize()
    y_dev = y.cpu().numpy().view(np.uint64).copy()
    hA, hx = A.cpu().numpy().view(np.uint64).copy(), x.cpu().numpy().view(np.uint64).copy()
    hy = np.zeros((m, 2), dtype=np.uint64)
    saved = Q.get_mode()
    Q.set_mode("reference")
    Q.qgemv("R", "N", m, m, 1.0, hA, m, hx, 1, 0.0, hy, 1)
    ts = []
    for _ in range(7):
        hy[:] = 0
        t0 = time.perf_counter()
        Q.qgemv("R", "N", m, m, 1.0, hA, m, hx, 1, 0.0, hy, 1)
        ts.append(time.perf_counter() - t0)
    Q.set_mode(saved)
    host_ms = statistics.median(ts) * 1e3
    del A, x, y
    torch.cuda.empty_cache()
    return {"workload": f"qgemv RowMajor {m}x{m}, alpha=1, beta=0, entries a + b 2^-60, REFERENCE order",
            "device": {"ms": round(ms, 4), "quad_GFLOPs": round(flops / (ms * 1e-3) / 1e9, 2)},
            "device_other_forms": other,
            "host_abi": {"ms": round(host_ms, 3), "quad_GFLOPs": round(flops / (host_ms * 1e-3) / 1e9, 2),
                         "pcie_bytes": 16.0 * (m * m + m + 2 * m)},
            "reference_published_GFLOPs": 1.6, "reference_published_source": "README.md:23 (SLEEF, OpenMP CPU)",
            "_check": (hA, hx, y_dev, hy)}


def bench_memory(Q, torch, dev, stream, n, m, steps):
    # a stream of its own (default priority): work on the null stream is ordered against every
    # other blocking stream of the process
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        return _bench_memory(Q, torch, dev, stream, n, m, steps)


def _bench_memory(Q, torch, dev, stream, n, m, steps):
    sh = stream.cuda_stream
    x = torch.empty((n, 2), dtype=torch.int64, device=dev)
    y = torch.empty((n, 2), dtype=torch.int64, device=dev)
    A = torch.empty((m * m, 2), dtype=torch.int64, device=dev)
    xv = torch.empty((m, 2), dtype=torch.int64, device=dev)
    yv0 = torch.empty((m, 2), dtype=torch.int64, device=dev)
    yv = torch.empty((m, 2), dtype=torch.int64, device=dev)
    res = torch.empty((1, 2), dtype=torch.int64, device=dev)
    for t, seed in ((x, 1), (y, 2), (A, 3), (xv, 4), (yv0, 5)):
        Q.dev_fill(Q.FILL_UNIFORM_FULL, seed, t.shape[0], t.data_ptr(), stream=sh)
    gemv_bytes = 16.0 * (m * m + m + 2 * m)
    out = {}
    for form in ("dd", "exact"):
        mode = Q.MODES[form]
        kernels = [("qdot", lambda: Q.dev_dot(mode, n, x.data_ptr(), 1, y.data_ptr(), 1, 1, res.data_ptr(), sh),
                    32.0 * n)]
        for lay, name in ((0, "qgemv_rowmajor"), (1, "qgemv_colmajor")):
            kernels.append((name, (lambda L: lambda: Q.dev_gemv(mode, L, m, m, 1.5, A.data_ptr(), m, xv.data_ptr(), 1,
                                                                 -0.5, yv.data_ptr(), 1, sh))(lay), gemv_bytes))
        for _name, f, _b in kernels:   # warm-up: ~100 calls (tens of ms) so clocks settle under load
            yv.copy_(yv0)
            for _ in range(100):
                f()
        torch.cuda.synchronize()
        res_f = {}
        for name, f, b in kernels:
            # back-to-back calls between two events (y keeps being updated: beta = -0.5 keeps it
            # bounded), so per-call launch gaps are amortised as in a stream of calls
            yv.copy_(yv0)
            torch.cuda.synchronize()
            s_ev, e_ev = events(torch)
            s_ev.record(stream)
            for _ in range(steps):
                f()
            e_ev.record(stream)
            torch.cuda.synchronize()
            ms = s_ev.elapsed_time(e_ev) / steps
            gbs = b / (ms * 1e-3) / 1e9
            res_f[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                           "algorithmic_bytes": b}
        out[form] = res_f
    return out


def bench_batched(Q, torch, dev, stream, steps, gemm_count, gemv_count, seed=77):
    """Config: 4096 independent qgemm problems with m, n, k uniform in [8, 256] and 16384 qgemv
    problems with m, n uniform in [32, 1024], RowMajor, entries and per-problem alpha, beta
    full-mantissa uniform [-1, 1) (device generator).  One grouped persistent launch per batch,
    tile configuration chosen by the autotuner, captured once in a graph and replayed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = {}
    sh = stream.cuda_stream

    def timed_graph(call):
        call()                                  # allocates the work-list buffer outside capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                call(s.cuda_stream)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        s_ev, e_ev = events(torch)
        s_ev.record(stream)
        for _ in range(steps):
            g.replay()
        e_ev.record(stream)
        torch.cuda.synchronize()
        return s_ev.elapsed_time(e_ev) / steps

    # ---- gemm batch
    dims = rng.integers(8, 257, size=(gemm_count, 3))
    sizes = dims[:, 0] * dims[:, 2] + dims[:, 2] * dims[:, 1] + dims[:, 0] * dims[:, 1]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    pool = torch.empty((int(offs[-1]), 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_UNIFORM_FULL, seed, pool.shape[0], pool.data_ptr(), stream=sh)
    ab = torch.empty((2 * gemm_count, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_UNIFORM_FULL, seed + 1, ab.shape[0], ab.data_ptr(), stream=sh)
    abh = ab.cpu().numpy().view(np.uint64)
    desc = np.zeros(gemm_count, Q.GEMM_PROBLEM)
    base = pool.data_ptr()
    for p in range(gemm_count):
        m, n, k = (int(v) for v in dims[p])
        o = int(offs[p])
        desc[p] = (m, n, k, k, n, n, base + 16 * o, base + 16 * (o + m * k), base + 16 * (o + m * k + k * n),
                   abh[2 * p], abh[2 * p + 1], 0)
    ddesc = torch.from_numpy(desc.view(np.uint8)).to(dev)
    flops = float(np.sum(2.0 * dims[:, 0] * dims[:, 1] * dims[:, 2]))
    g = {"problems": gemm_count, "dims": "m, n, k uniform in [8, 256]", "quad_flops": flops}
    for form in ("dd", "reference"):
        mode = Q.MODES[form]
        cfg = Q.dev_gemm_batched_tune(mode, 0, gemm_count, ddesc.data_ptr(), sh)
        ms = timed_graph(lambda s=sh: Q.dev_gemm_batched(mode, 0, gemm_count, ddesc.data_ptr(), s))
        g[form] = {"ms": round(ms, 3), "problems_per_s": round(gemm_count / (ms * 1e-3), 1),
                   "quad_GFLOPs": round(flops / (ms * 1e-3) / 1e9, 1), "tuned_config": Q.gemm_batched_config(mode, cfg)}
    # OZAKI form (slice count tuned on the batch's own operands): the eager call in timed_graph
    # sizes the workspace (one read-back); the captured call reuses those totals and every replay
    # checks them on the device
    mode = Q.MODES["ozaki"]
    cfg = Q.dev_gemm_batched_tune(mode, 0, gemm_count, ddesc.data_ptr(), sh)
    ms = timed_graph(lambda s=sh: Q.dev_gemm_batched(mode, 0, gemm_count, ddesc.data_ptr(), s))
    status = Q.dev_exact_status()
    g["ozaki"] = {"ms": round(ms, 3), "problems_per_s": round(gemm_count / (ms * 1e-3), 1),
                  "quad_GFLOPs": round(flops / (ms * 1e-3) / 1e9, 1), "tuned_config": Q.gemm_batched_config(mode, cfg),
                  "repaired_entries": int(Q.dev_ozaki_repairs()), "graph": True, "status": int(status)}
    out["qgemm"] = g
    del pool, ab, ddesc
    torch.cuda.empty_cache()

    # ---- gemv batch
    mn = rng.integers(32, 1025, size=(gemv_count, 2))
    asz = mn[:, 0] * mn[:, 1]
    offs = np.concatenate([[0], np.cumsum(asz + mn[:, 1] + mn[:, 0])])
    pool = torch.empty((int(offs[-1]), 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_UNIFORM_FULL, seed + 2, pool.shape[0], pool.data_ptr(), stream=sh)
    ab = torch.empty((2 * gemv_count, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_UNIFORM_FULL, seed + 3, ab.shape[0], ab.data_ptr(), stream=sh)
    abh = ab.cpu().numpy().view(np.uint64)
    desc = np.zeros(gemv_count, Q.GEMV_PROBLEM)
    base = pool.data_ptr()
    for p in range(gemv_count):
        m, n = int(mn[p, 0]), int(mn[p, 1])
        o = int(offs[p])
        desc[p] = (m, n, n, 1, 1, base + 16 * o, base + 16 * (o + m * n), base + 16 * (o + m * n + n),
                   abh[2 * p], abh[2 * p + 1])
    ddesc = torch.from_numpy(desc.view(np.uint8)).to(dev)
    nbytes = float(np.sum(16.0 * (asz + mn[:, 1] + 2 * mn[:, 0])))
    v = {"problems": gemv_count, "dims": "m, n uniform in [32, 1024]", "algorithmic_bytes": nbytes}
    for form in ("dd", "reference"):
        mode = Q.MODES[form]
        ms = timed_graph(lambda s=sh: Q.dev_gemv_batched(mode, 0, gemv_count, ddesc.data_ptr(), s))
        v[form] = {"ms": round(ms, 3), "problems_per_s": round(gemv_count / (ms * 1e-3), 1),
                   "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    out["qgemv"] = v
    del pool, ab, ddesc
    torch.cuda.empty_cache()
    return out


def bench_host_abi(Q, torch, dev, N, m, n, reps):
    """The reference's C ABI with host buffers (quadblas_qgemm in the OZAKI form, quadblas_qgemv
    RowMajor and quadblas_qdot in the DD form, same workloads as above): each call copies its
    operands to HBM and its result back over PCIe and returns when the result is in host memory,
    as a drop-in caller's call does.  Wall-clock per call (median of `reps` after one warm-up),
    with pageable host buffers (what a caller's malloc gives) and with pinned ones (hipHostMalloc,
    torch's pin_memory).  pcie_bytes: the operands in plus the result out."""
    import numpy as np
    sh = torch.cuda.current_stream().cuda_stream
    gA = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    gB = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 11, N * N, gA.data_ptr(), -30, 30, stream=sh)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 12, N * N, gB.data_ptr(), -30, 30, stream=sh)
    gM = torch.empty((m * m, 2), dtype=torch.int64, device=dev)
    gv = torch.empty((m, 2), dtype=torch.int64, device=dev)
    gx = torch.empty((n, 2), dtype=torch.int64, device=dev)
    gy = torch.empty((n, 2), dtype=torch.int64, device=dev)
    for t, seed in ((gM, 3), (gv, 4), (gx, 1), (gy, 2)):
        Q.dev_fill(Q.FILL_UNIFORM_FULL, seed, t.shape[0], t.data_ptr(), stream=sh)
    torch.cuda.synchronize()
    saved = Q.get_mode()

    def timed(f):
        f()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts) * 1e3

    out = {}
    for pinned in (False, True):
        def host(t):
            h = torch.empty(t.shape, dtype=t.dtype, pin_memory=pinned)
            h.copy_(t)
            return h
        hA, hB, hM, hv, hx, hy = (host(t) for t in (gA, gB, gM, gv, gx, gy))
        hC = torch.zeros((N * N, 2), dtype=torch.int64, pin_memory=pinned)
        hw = torch.zeros((m, 2), dtype=torch.int64, pin_memory=pinned)
        a, b, c, M_, v, w, x, y = (t.numpy().view(np.uint64) for t in (hA, hB, hC, hM, hv, hw, hx, hy))
        r = {}
        Q.set_mode("ozaki")
        ms = timed(lambda: Q.qgemm("R", "N", "N", N, N, N, 1.0, a, N, b, N, 0.0, c, N))
        pb = 16.0 * 4 * N * N   # A, B, C in; C out
        r["qgemm"] = {"ms": round(ms, 2), "quad_GFLOPs": round(2.0 * N ** 3 / (ms * 1e-3) / 1e9, 1),
                      "pcie_bytes": pb, "GB/s": round(pb / (ms * 1e-3) / 1e9, 1)}
        Q.set_mode("dd")
        ms = timed(lambda: Q.qgemv("R", "N", m, m, 1.5, M_, m, v, 1, -0.5, w, 1))
        pb = 16.0 * (m * m + m + 2 * m)
        r["qgemv_rowmajor"] = {"ms": round(ms, 2), "pcie_bytes": pb, "GB/s": round(pb / (ms * 1e-3) / 1e9, 1)}
        ms = timed(lambda: Q.qdot(n, x, 1, y, 1))
        pb = 32.0 * n
        r["qdot"] = {"ms": round(ms, 2), "pcie_bytes": pb, "GB/s": round(pb / (ms * 1e-3) / 1e9, 1)}
        out["pinned" if pinned else "pageable"] = r
        del hA, hB, hC, hM, hv, hw, hx, hy, a, b, c, M_, v, w, x, y
    Q.set_mode(saved)
    del gA, gB, gM, gv, gx, gy
    torch.cuda.empty_cache()
    return out


def bench_ill_conditioned(Q, torch, dev, M, N, K, plain=True):
    """qgemm M x N x K ColMajor, alpha = 1, beta = 0: rows of A carry +-1e22 u pairs along k and
    the columns of B equal values on each pair (qblas_hip_fill_2d kinds 4 and 3), so the exact
    entries are ~1e-20 of the largest products.  OZAKI form with the library's automatic choice of
    method and precision (the probe asks for more than the CRT method's 158 bits: the slice
    method), the DD form and the plain form (one call, no warm-up), one timed call each, on a stream
    of its own so that the OZAKI call's whole workspace is new (its size is what the call adds to
    HBM in use).  Only the OZAKI form meets 2^-104 here (tests/test_gpu_ozaki_crt.py measures all
    three on sampled entries against the exact value)."""
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    A = torch.empty((M * K, 2), dtype=torch.int64, device=dev)
    B = torch.empty((K * N, 2), dtype=torch.int64, device=dev)
    C = torch.zeros((M * N, 2), dtype=torch.int64, device=dev)
    Q.dev_fill_2d(4, 5, M, K, M, 1, A.data_ptr(), stream=sh)
    Q.dev_fill_2d(3, 5, N, K, K, 0, B.data_ptr(), stream=sh)
    torch.cuda.synchronize()
    flops = 2.0 * M * N * K
    operand_bytes = 16.0 * (M * K + K * N + M * N)
    out = {"workload": f"qgemm {M}x{N}x{K} ColMajor alpha=1 beta=0, entries ~1e-20 of the largest products "
                       f"(+-1e22 u pairs)", "quad_flops": flops, "operand_bytes": operand_bytes}
    for form in ("ozaki", "dd") + (("reference",) if plain else ()):
        mode = Q.MODES[form]
        if form == "ozaki":
            Q.set_phase_timing(True)
        call = lambda: Q.dev_gemm(mode, 1, M, N, K, 1.0, A.data_ptr(), M, B.data_ptr(), K, 0.0, C.data_ptr(), M, sh)
        free0, total = torch.cuda.mem_get_info()
        if form != "reference":
            call()
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        s_ev, e_ev = events(torch)
        s_ev.record(stream)
        call()
        e_ev.record(stream)
        torch.cuda.synchronize()
        ms = s_ev.elapsed_time(e_ev)
        r = {"ms": round(ms, 2), "quad_GFLOPs": round(flops / (ms * 1e-3) / 1e9, 1),
             "hbm_GBs_algorithmic": round(operand_bytes / (ms * 1e-3) / 1e9, 1)}
        if form != "reference":
            ws = max(0, free0 - free1)
            r["workspace_GB"] = round(ws / 1e9, 2)
            r["hbm_call_GB"] = round((operand_bytes + ws) / 1e9, 2)
        if form == "ozaki":
            S = Q.ozaki_last_slices()
            ph = Q.phase_times()
            Q.set_phase_timing(False)
            r["method"] = "crt" if Q.ozaki_last_method() == Q.OZAKI_CRT else "slices"
            r["slices"] = S
            r["repaired_entries"] = int(Q.dev_ozaki_repairs(sh))
            r["phases_ms"] = {k: round(v, 3) for k, v in ph.items()}
            ops = 2.0 * M * N * K * (S * (S + 1) // 2)
            r["mfma"] = {"int8_ops": ops, "TOP/s": round(ops / (ph["products"] * 1e-3) / 1e12, 1),
                         "frac_of_peak": round(ops / (ph["products"] * 1e-3) / 1e12 / INT8_PEAK_TOPS, 4)}
        out["plain" if form == "reference" else form] = r
    del A, B, C
    torch.cuda.empty_cache()
    return out


def summarize_pmc(fetch_csv, write_csv):
    """HBM bytes per launch of the OZAKI kernels from two rocprofv3 counter-collection CSVs.
    Each counter is scaled by what it reads for the 1 GiB calibration copy (`__amd_rocclr_copyBuffer`
    reads and writes exactly 2^30 bytes).  Kernels are keyed by name in launch order; a name seen more
    than once gets #2, #3 ... (the slice-count probe runs the scale kernels before the call's own)."""
    import csv

    def read(path, counter):
        rows = []
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
        rows.sort()
        return rows

    def short(name):
        base = name.split("(")[0].replace("void ", "").replace("qbk::", "").strip()
        return base

    fetch, write = read(fetch_csv, "FETCH_SIZE"), read(write_csv, "WRITE_SIZE")
    copy_f = [v for _, n, v in fetch if "copyBuffer" in n]
    copy_w = [v for _, n, v in write if "copyBuffer" in n]
    if not copy_f or not copy_w:
        raise SystemExit("no calibration copy in the counter files (run with --copy-calibration)")
    gib = float(1 << 30)
    fs, ws = gib / (copy_f[0] * 1024), gib / (copy_w[0] * 1024)
    out = {"calibration": {"copy_bytes": 1 << 30, "copy_fetch_kib": copy_f[0], "copy_write_kib": copy_w[0],
                           "fetch_scale": round(fs, 4), "write_scale": round(ws, 4)},
           "note": "HBM bytes per launch = counter KiB * 1024 * scale"}
    wmap = {d: v for d, _, v in write}
    seen = {}
    for d, name, fv in fetch:
        if "k_oz_" not in name and "k_crt_" not in name:
            continue
        key = short(name)
        seen[key] = seen.get(key, 0) + 1
        if seen[key] > 1:
            key = f"{key}#{seen[key]}"
        wv = wmap.get(d, 0.0)
        rb, wb = int(fv * 1024 * fs), int(wv * 1024 * ws)
        out[key] = {"fetch_kib": fv, "write_kib": wv, "read_bytes": rb, "write_bytes": wb,
                    "hbm_bytes_per_launch": rb + wb}
    for prefix in ("k_oz_gemm", "k_crt_gemm"):
        gemm = [k for k in out if k.startswith(prefix) and k != prefix]
        if gemm:
            out[prefix] = dict(out[gemm[-1]], kernel=gemm[-1])
    return out


def load_traffic():
    path = os.path.join(ROOT, "profiles", "ozaki_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_info():
    """Host CPU model and the threads the OpenMP restatements use (OMP_NUM_THREADS, else all)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    return model, threads


def cpu_baseline_gemm(N, rows, fmaq_rows):
    """The oracle's reference-order qgemm (level3.hpp blocking, kc = 126, soft binary128 FMA, OpenMP
    over rows x columns) on rows x N of C for the same N^3 problem's distribution; also with
    libquadmath's fmaq as the FMA (a smaller slab)."""
    import oracle as O
    model, threads = cpu_info()
    B = O.fill(1, 12, N * N, -30, 30)
    out = None
    for variant, r in (("oracle soft-binary128 FMA", rows), ("libquadmath fmaq", fmaq_rows)):
        A = O.fill(1, 11, r * N, -30, 30)
        C = O.fill(0, 13, r * N)
        C[:] = 0
        t0 = time.perf_counter()
        if variant.startswith("libquadmath"):
            with O.use_fmaq():
                O.gemm(0, r, N, N, 1.0, A, N, B, N, 0.0, C, N)
        else:
            O.gemm(0, r, N, N, 1.0, A, N, B, N, 0.0, C, N)
        t = time.perf_counter() - t0
        v = {"value": round(2.0 * r * N * N / t / 1e9, 4), "unit": "quad-equiv GFLOP/s", "cores": threads,
             "kind": "port", "cpu_model": model,
             "sample": f"reference-order qgemm on a {r} x {N} slab of C ({r}x{N}x{N} FMAs, same entry distribution), "
                       f"{variant}, OpenMP; {t:.1f} s, extrapolated to the full problem: {t * N / r:.0f} s",
             "seconds": round(t, 2), "extrapolated_full_problem_s": round(t * N / r, 1)}
        if out is None:
            out = v
        else:
            out["fmaq_variant"] = {k: v[k] for k in ("value", "sample", "seconds", "extrapolated_full_problem_s")}
    return out


def cpu_baseline_memory(n, m):
    """The oracle's reference-order qdot (T chunks) + qgemv RowMajor + ColMajor, one step."""
    import oracle as O
    model, T = cpu_info()
    x = O.fill(0, 1, n)
    y = O.fill(0, 2, n)
    A = O.fill(0, 3, m * m)
    xv = O.fill(0, 4, m)
    yv0 = O.fill(0, 5, m)
    t0 = time.perf_counter()
    O.dot(n, x, 1, y, 1, T)
    yv = yv0.copy()
    O.gemv(0, m, m, 1.5, A, m, xv, 1, -0.5, yv, 1)
    yv = yv0.copy()
    O.gemv(1, m, m, 1.5, A, m, xv, 1, -0.5, yv, 1)
    t = time.perf_counter() - t0
    step_bytes = 32.0 * n + 2 * 16.0 * (m * m + m + 2 * m)
    return {"value": round(step_bytes / t / 1e9, 3), "unit": "GB/s", "cores": T, "kind": "port", "cpu_model": model,
            "sample": "one step: qdot 2^24 + qgemv 8192^2 RowMajor + ColMajor, reference order, oracle soft "
                      "binary128 FMA, OpenMP", "seconds": round(t, 3)}


def cpu_config1(c1):
    """The config-1 qgemv on the host cores: the oracle's restatement of level2.hpp:15-50 (OpenMP over
    rows), two warm-ups then the median of 7 runs (the reference's perf_compare.cpp:32-46 protocol),
    with the oracle's FMA and with libquadmath fmaq; and the check that the GPU result (device call
    and C ABI call) equals the restatement bit for bit."""
    import numpy as np
    import oracle as O
    model, threads = cpu_info()
    hA, hx, y_dev, y_host = c1.pop("_check")
    m = hx.shape[0]
    res = {}
    for variant in ("oracle", "fmaq"):
        ts = []
        want = None
        for rep in range(9):
            y = np.zeros((m, 2), dtype=np.uint64)
            t0 = time.perf_counter()
            if variant == "fmaq":
                with O.use_fmaq():
                    O.gemv(0, m, m, 1.0, hA, m, hx, 1, 0.0, y, 1)
            else:
                O.gemv(0, m, m, 1.0, hA, m, hx, 1, 0.0, y, 1)
            if rep >= 2:
                ts.append(time.perf_counter() - t0)
            want = y
        ms = statistics.median(ts) * 1e3
        res[variant] = {"ms": round(ms, 3), "quad_GFLOPs": round(2.0 * m * m / (ms * 1e-3) / 1e9, 3), "y": want}
    want = res["oracle"].pop("y")
    res["fmaq"].pop("y")
    c1["cpu_baseline"] = {"value": res["oracle"]["quad_GFLOPs"], "unit": "quad-equiv GFLOP/s", "cores": threads,
                          "kind": "port", "cpu_model": model,
                          "sample": f"the whole qgemv {m}x{m}, median of 7 after 2 warm-ups, OpenMP over rows",
                          "ms": res["oracle"]["ms"], "fmaq_variant": res["fmaq"]}
    c1["bit_exact_vs_restatement"] = {"device": bool(np.array_equal(y_dev, want)),
                                      "host_abi": bool(np.array_equal(y_host, want))}
    return c1


def q_fraction(w):
    """Exact value of a finite binary128 given as (lo, hi) uint64 words."""
    from fractions import Fraction
    lo, hi = int(w[0]), int(w[1])
    be, f = (hi >> 48) & 0x7FFF, ((hi & ((1 << 48) - 1)) << 64) | lo
    v = Fraction(f if be == 0 else f | (1 << 112)) * Fraction(2) ** ((1 if be == 0 else be) - 16383 - 112)
    return -v if hi >> 63 else v


def cpu_error_checks(N, forms):
    """Error statement of every qgemm form on the fixed 4096 sampled entries: relative error
    against the exact value (the oracle's exact accumulator, rounded once) and against the
    REFERENCE form's (plain) result at the same entries; the plain form's samples are also checked
    bit for bit against the oracle's per-entry restatement of the reference's blocked order
    (level3.hpp:215-336, kc = 126: a 1 x 1 x N problem takes the same path per entry)."""
    import numpy as np
    import oracle as O
    ii, jj, _ = sample_index(N)
    A = O.fill(1, 11, N * N, -30, 30)
    B = O.fill(1, 12, N * N, -30, 30)
    ref = forms.get("plain", {}).get("samples")
    for name, f in forms.items():
        samp = f.pop("samples", None)
        if samp is None:
            continue
        _, rel = O.exact_gemm_samples(ii, jj, N, 1.0, A, (N, 1), B, (N, 1), 0.0, None, (N, 1), cand=samp)
        e = {"samples": int(len(ii)), "max_rel_vs_exact": float(rel.max()), "median_rel_vs_exact": float(np.median(rel)),
             "log2_max_vs_exact": round(float(np.log2(max(rel.max(), 1e-300))), 2)}
        if ref is not None:
            rr = []
            for a, b in zip(samp, ref):
                fa, fb = q_fraction(a), q_fraction(b)
                rr.append(float(abs(fa - fb) / abs(fb)) if fb != 0 else (0.0 if fa == 0 else float("inf")))
            e["max_rel_vs_reference"] = max(rr)
            e["median_rel_vs_reference"] = float(np.median(rr))
        f["error"] = e
    if ref is not None:
        mism = 0
        c = np.zeros((1, 2), dtype=np.uint64)
        for s_, (i, j) in enumerate(zip(ii, jj)):
            c[:] = 0
            O.gemm(0, 1, 1, N, 1.0, np.ascontiguousarray(A[i * N:(i + 1) * N]), N,
                   np.ascontiguousarray(B[j::N]), 1, 0.0, c, 1)
            mism += int(not np.array_equal(c[0], ref[s_]))
        forms["plain"]["error"]["restatement_mismatches"] = mism


def cpu_baseline_batched(gemm_count, gemv_count, sample_gemm, sample_gemv, seed=77):
    """A serial loop over the first problems of the same batches (the reference has no batched
    routine: one restatement call per problem, OpenMP inside, not across problems -- the reference's
    dot breaks inside an outer parallel region), extrapolated to the whole batch by its flops or bytes."""
    import numpy as np
    import oracle as O
    model, threads = cpu_info()
    rng = np.random.default_rng(seed)
    dims = rng.integers(8, 257, size=(gemm_count, 3))
    mn = rng.integers(32, 1025, size=(gemv_count, 2))
    out = {}
    # gemm batch: operands of problem p at offs[p] in the seed's kind-0 sequence (A | B | C)
    sizes = dims[:, 0] * dims[:, 2] + dims[:, 2] * dims[:, 1] + dims[:, 0] * dims[:, 1]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    P = min(sample_gemm, gemm_count)
    pool = O.fill(0, seed, int(offs[P]))
    ab = O.fill(0, seed + 1, 2 * P)
    t0 = time.perf_counter()
    for p in range(P):
        m, n, k = (int(v) for v in dims[p])
        o = int(offs[p])
        Cp = pool[o + m * k + k * n:o + m * k + k * n + m * n]
        O.gemm(0, m, n, k, ab[2 * p], pool[o:o + m * k], k, pool[o + m * k:o + m * k + k * n], n, ab[2 * p + 1], Cp, n)
    t = time.perf_counter() - t0
    fl_s = float(np.sum(2.0 * dims[:P, 0] * dims[:P, 1] * dims[:P, 2]))
    fl = float(np.sum(2.0 * dims[:, 0] * dims[:, 1] * dims[:, 2]))
    out["qgemm_reference"] = {"value": round(gemm_count / (t * fl / fl_s), 1), "unit": "problems/s",
                              "quad_GFLOPs": round(fl_s / t / 1e9, 3), "cores": threads, "kind": "port",
                              "cpu_model": model, "seconds": round(t, 2),
                              "sample": f"serial loop over the first {P} of {gemm_count} problems (reference order), "
                                        f"extrapolated by flops"}
    del pool
    asz = mn[:, 0] * mn[:, 1]
    offs = np.concatenate([[0], np.cumsum(asz + mn[:, 1] + mn[:, 0])])
    P = min(sample_gemv, gemv_count)
    pool = O.fill(0, seed + 2, int(offs[P]))
    ab = O.fill(0, seed + 3, 2 * P)
    t0 = time.perf_counter()
    for p in range(P):
        m, n = int(mn[p, 0]), int(mn[p, 1])
        o = int(offs[p])
        y = pool[o + m * n + n:o + m * n + n + m].copy()
        O.gemv(0, m, n, ab[2 * p], pool[o:o + m * n], n, pool[o + m * n:o + m * n + n], 1, ab[2 * p + 1], y, 1)
    t = time.perf_counter() - t0
    by_s = float(np.sum(16.0 * (asz[:P] + mn[:P, 1] + 2 * mn[:P, 0])))
    by = float(np.sum(16.0 * (asz + mn[:, 1] + 2 * mn[:, 0])))
    out["qgemv_reference"] = {"value": round(gemv_count / (t * by / by_s), 1), "unit": "problems/s",
                              "GB/s": round(by_s / t / 1e9, 3), "cores": threads, "kind": "port", "cpu_model": model,
                              "seconds": round(t, 2),
                              "sample": f"serial loop over the first {P} of {gemv_count} problems (reference order), "
                                        f"extrapolated by bytes"}
    return out


def cpu_baseline_ill(M, N, K, rows, cols):
    """The ill-conditioned qgemm's restatement (the layout-correct ColMajor blocked order; the
    reference's own ColMajor blocked path writes C row-major) on a rows x cols block of C,
    extrapolated to M x N."""
    import numpy as np
    import oracle as O
    model, threads = cpu_info()
    Asub = np.zeros((rows * K, 2), dtype=np.uint64)
    for r in range(rows):                       # A(r, c) = value number r K + c, stored at r + c rows
        Asub[r::rows] = O.fill(4, 5, K, start=r * K)
    Bsub = O.fill_pair_y(5, cols * K)           # B(l, j) = value number j K + l at l + j K
    C = np.zeros((rows * cols, 2), dtype=np.uint64)
    t0 = time.perf_counter()
    O.gemm(1, rows, cols, K, 1.0, Asub, rows, Bsub, K, 0.0, C, rows)
    t = time.perf_counter() - t0
    scale = (M * N) / (rows * cols)
    return {"value": round(2.0 * rows * cols * K / t / 1e9, 4), "unit": "quad-equiv GFLOP/s", "cores": threads,
            "kind": "port", "cpu_model": model, "seconds": round(t, 2), "extrapolated_full_problem_s": round(t * scale, 1),
            "sample": f"reference-order (layout-correct ColMajor) qgemm on a {rows} x {cols} block of C "
                      f"({rows}x{cols}x{K} FMAs), oracle soft FMA, OpenMP, extrapolated x{scale:.0f}"}


def main():
    args = parse()
    if args.summarize_pmc:
        print(json.dumps(summarize_pmc(*args.summarize_pmc), indent=1))
        sys.exit(0)
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import qblas_amd as Q

    Q.lib()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    N = args.size
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # the memory-bound routines run first, not right after seconds of full-chip integer work
    # (the plain qgemm), so that they are timed on a chip in the same state as when measured alone
    memory = None
    if not args.no_memory:
        memory = bench_memory(Q, torch, dev, stream, args.n_dot, args.m, max(args.steps, 50))
    methods = {"auto": Q.OZAKI_AUTO, "slices": Q.OZAKI_SLICES, "crt": Q.OZAKI_CRT}
    oz = bench_ozaki(Q, torch, dev, stream, N, args.steps, args.warmup, args.slices, methods[args.method],
                     args.crt_bits)
    S = oz["slices"]
    if oz["method"] == "crt":
        ops = 2.0 * N * N * N * oz["moduli"]
        kname, kdesc = "k_crt_gemm", f"k_crt_gemm ({oz['moduli']} int8 GEMMs, one per modulus, " \
                                      f"v_mfma_i32_32x32x32_i8)"
    else:
        ops = 2.0 * N * N * N * (S * (S + 1) // 2)
        kname, kdesc = "k_oz_gemm", f"k_oz_gemm<{S}> (slice products, v_mfma_i32_32x32x32_i8)"
    prod_ms = oz["phases_ms"]["products"]
    achieved = ops / (prod_ms * 1e-3) / 1e12
    traffic = load_traffic()
    roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                "frac": round(achieved / INT8_PEAK_TOPS, 4),
                "traffic": (traffic or {}).get(kname, {}).get("hbm_bytes_per_launch"),
                "kernel": kdesc, "algorithmic_ops_per_launch": ops, "launch_ms": prod_ms}
    forms = {"ozaki": ozaki_form(oz)}
    if oz["method"] == "crt" and not args.no_slices_form:
        forms["ozaki_slices"] = ozaki_form(bench_ozaki(Q, torch, dev, stream, N, max(args.steps // 2, 1), 1, 0,
                                                       Q.OZAKI_SLICES, 0))
    forms["dd"] = bench_form_once(Q, torch, dev, stream, N, Q.MODE_DD)
    forms["dd"]["speedup_vs_plain"] = None
    macs = float(N) ** 3
    fp64_tflops = macs * DD_GEMM_FP64_FLOPS_PER_MAC / (forms["dd"]["ms"] * 1e-3) / 1e12
    forms["dd"]["fp64"] = {"flops_per_mac": DD_GEMM_FP64_FLOPS_PER_MAC,
                           "instructions_per_mac": DD_GEMM_FP64_INSTS_PER_MAC,
                           "TFLOP/s": round(fp64_tflops, 1), "frac_of_peak": round(fp64_tflops / FP64_PEAK_TFLOPS, 4),
                           "issue_frac": round(macs * DD_GEMM_FP64_INSTS_PER_MAC / (forms["dd"]["ms"] * 1e-3) /
                                               (FP64_PEAK_TFLOPS / 2 * 1e12), 4)}
    if not args.no_plain:
        forms["plain"] = bench_plain(Q, torch, dev, stream, N)
        forms["dd"]["speedup_vs_plain"] = round(forms["plain"]["ms"] / forms["dd"]["ms"], 1)
        for key in ("ozaki", "ozaki_slices"):
            if key in forms:
                forms[key]["speedup_vs_plain"] = round(forms["plain"]["ms"] / forms[key]["ms"], 1)
    torch.cuda.empty_cache()
    config1 = bench_config1(Q, torch, dev, stream, 1500, 200)
    # the host-buffer calls run before the legs that allocate tens of GB: on a box where 74 GB had
    # been allocated and freed, the copy of C in no longer overlapped the qgemm's kernels
    host_abi = None
    if not args.no_host:
        host_abi = bench_host_abi(Q, torch, dev, N, args.m, args.n_dot, 3)
    ill = None
    if not args.no_ill:
        ill = bench_ill_conditioned(Q, torch, dev, 8192, 8192, 16384, plain=not args.no_ill_plain)
    batched = None
    if not args.no_batched:
        batched = bench_batched(Q, torch, dev, stream, max(args.steps, 3), args.batch_gemm, args.batch_gemv)
    cpu = cpu_mem = None
    if not args.no_cpu:
        # the CPU legs: baselines on the host cores, and the oracle as the checker of the GPU
        # results (sampled error of every qgemm form, config 1 bit for bit)
        cpu_error_checks(N, forms)
        cpu = cpu_baseline_gemm(N, args.cpu_rows, args.cpu_rows // 16)
        cpu_mem = cpu_baseline_memory(args.n_dot, args.m)
        cpu_config1(config1)
        if batched is not None:
            batched["cpu_baseline"] = cpu_baseline_batched(args.batch_gemm, args.batch_gemv, 256, 256)
        if ill is not None:
            ill["cpu_baseline"] = cpu_baseline_ill(8192, 8192, 16384, 64, 1024)
    else:
        config1.pop("_check", None)
    for f in forms.values():
        f.pop("samples", None)

    line = {
        "metric": METRIC, "value": round(oz["quad_GFLOPs"], 2), "unit": "quad-equiv GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(oz["ms"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i8 (int8 x int8 -> int32 MFMA; binary128 in/out)",
        "data": "synthetic (seeded device generator: random sign, log-uniform magnitude [2^-30, 2^30), "
                "full 113-bit mantissa)",
        "config": {"workload": f"qgemm {N}x{N}x{N} RowMajor alpha=1 beta=0",
                   "form": f"ozaki, {oz['method']} method" + (f" ({oz['crt_bits']} bits, {oz['moduli']} moduli)"
                                                              if oz["method"] == "crt" else f" ({S} slices)"),
                   "error_bound": "each entry <= 2^-105.9 rel (certified) or correctly rounded (repaired)"},
        "roofline": roofline, "cpu_baseline": cpu, "qgemm_forms": forms, "memory_bound": memory,
        "cpu_baseline_memory_bound": cpu_mem, "config1_qgemv_1500": config1, "batched": batched,
        "ill_conditioned": ill, "host_abi": host_abi, "ozaki_status": oz["status"], "wall_s": round(oz["wall_s"], 3),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
