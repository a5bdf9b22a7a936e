"""Benchmark of the QuadBLAS binary128 hot path on one MI355X.

Workload (the reference's memory-bound Level 1/2 configuration, synthetic seeded inputs generated
on the device: every entry is a + b*2^-60 with a, b uniform [-1, 1) doubles, summed in binary128):
  qdot  n = 2^24                          (32 B per element read)
  qgemv 8192 x 8192 RowMajor, alpha = 1.5, beta = -0.5
  qgemv 8192 x 8192 ColMajor, alpha = 1.5, beta = -0.5
One step = one call of each.  value = algorithmic HBM bytes of the step / step time (GB/s), with
all operands resident in HBM before the timed region.  Algorithmic bytes: dot 2*16*n, gemv
16*(m*n + n + 2*m) (A and x read, y read and written).

Also reported: per-kernel times, the roofline object for the dominant kernel, and the CPU
baseline (the oracle's restatement of the reference's OpenMP loops on this host's cores).
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form reference|exact|dd|auto]
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
METRIC = "quad-equiv GFLOP/s, qgemm 4096³ at ≤2^-104 rel err; qdot/qgemv HBM GB/s"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--form", default="auto", choices=["auto", "reference", "exact", "dd"])
    p.add_argument("--n-dot", type=int, default=1 << 24)
    p.add_argument("--m", type=int, default=8192)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0 = OMP default)")
    p.add_argument("--gemm", default="", help="also time qgemm 4096^3 in these forms (comma list)")
    p.add_argument("--gemm-size", type=int, default=4096)
    return p.parse_args()


def pick_form(Q, form):
    if form != "auto":
        return Q.MODES[form]
    return Q.MODE_REFERENCE


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import qblas_amd as Q

    Q.lib()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    mode = pick_form(Q, args.form)
    form_name = [k for k, v in Q.MODES.items() if v == mode][0]

    n, m = args.n_dot, args.m
    T = Q.get_num_threads()
    x = torch.empty((n, 2), dtype=torch.int64, device=dev)
    y = torch.empty((n, 2), dtype=torch.int64, device=dev)
    A = torch.empty((m * m, 2), dtype=torch.int64, device=dev)
    xv = torch.empty((m, 2), dtype=torch.int64, device=dev)
    yv0 = torch.empty((m, 2), dtype=torch.int64, device=dev)
    yv = torch.empty((m, 2), dtype=torch.int64, device=dev)
    res = torch.empty((1, 2), dtype=torch.int64, device=dev)
    for t, seed in ((x, 1), (y, 2), (A, 3), (xv, 4), (yv0, 5)):
        Q.dev_fill(Q.FILL_UNIFORM_FULL, seed, t.shape[0], t.data_ptr(), stream=sh)
    torch.cuda.synchronize()

    def k_dot():
        Q.dev_dot(mode, n, x.data_ptr(), 1, y.data_ptr(), 1, T, res.data_ptr(), sh)

    def k_gemv(layout):
        def f():
            Q.dev_gemv(mode, layout, m, m, 1.5, A.data_ptr(), m, xv.data_ptr(), 1, -0.5, yv.data_ptr(), 1, sh)
        return f

    kernels = [("qdot", k_dot, 32.0 * n), ("qgemv_rowmajor", k_gemv(0), 16.0 * (m * m + m + 2 * m)),
               ("qgemv_colmajor", k_gemv(1), 16.0 * (m * m + m + 2 * m))]

    def reset_y():
        yv.copy_(yv0)

    for _ in range(args.warmup):
        for name, f, _b in kernels:
            reset_y()
            f()
    torch.cuda.synchronize()

    # timed region: K steps, per-kernel events on the launch stream
    ev = {name: [] for name, _, _ in kernels}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for name, f, _b in kernels:
            reset_y()
            s_ev, e_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s_ev.record(stream)
            f()
            e_ev.record(stream)
            ev[name].append((s_ev, e_ev))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_kernel_ms = {name: statistics.mean(s.elapsed_time(e) for s, e in ev[name]) for name in ev}
    step_ms = sum(per_kernel_ms.values())
    step_bytes = sum(b for _, _, b in kernels)
    value = step_bytes / (step_ms * 1e-3) / 1e9

    dom = max(kernels, key=lambda k: per_kernel_ms[k[0]])
    achieved = dom[2] / (per_kernel_ms[dom[0]] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom[0],
                "algorithmic_bytes": dom[2]}
    per_kernel = {name: {"ms": round(per_kernel_ms[name], 4),
                         "GB/s": round(b / (per_kernel_ms[name] * 1e-3) / 1e9, 1),
                         "hbm_frac": round(b / (per_kernel_ms[name] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                  for name, _, b in kernels}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, m, T, args.cpu_threads)

    gemm = {}
    for gform in [f for f in args.gemm.split(",") if f]:
        gemm[gform] = bench_gemm(Q, torch, dev, stream, Q.MODES[gform], args.gemm_size)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(step_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "binary128 (soft, u64/u32 integer ops)" if mode in (0, 1) else "binary128 via f64 double-double",
        "data": "synthetic (seeded device generator, full 113-bit mantissa)",
        "config": {"workload": "qdot n=2^24 + qgemv 8192x8192 RowMajor + ColMajor, alpha=1.5 beta=-0.5",
                   "form": form_name, "dot_T": T},
        "per_kernel": per_kernel, "roofline": roofline, "cpu_baseline": cpu,
        "wall_s": round(wall, 3),
    }
    if gemm:
        line["qgemm"] = gemm
    print(json.dumps(line))


def bench_gemm(Q, torch, dev, stream, mode, N, reps=2):
    """qgemm N^3 RowMajor, alpha = 1, beta = 0, entries with random sign, exponent uniform in
    [-30, 30) and a full 113-bit mantissa.  Returns time and quad-equiv GFLOP/s (2 N^3 / t)."""
    sh = stream.cuda_stream
    A = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    B = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    C = torch.empty((N * N, 2), dtype=torch.int64, device=dev)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 11, N * N, A.data_ptr(), -30, 30, stream=sh)
    Q.dev_fill(Q.FILL_LOG_UNIFORM, 12, N * N, B.data_ptr(), -30, 30, stream=sh)
    C.zero_()
    Q.dev_gemm(mode, 0, N, N, N, 1.0, A.data_ptr(), N, B.data_ptr(), N, 0.0, C.data_ptr(), N, sh)  # warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        s_ev, e_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s_ev.record(stream)
        Q.dev_gemm(mode, 0, N, N, N, 1.0, A.data_ptr(), N, B.data_ptr(), N, 0.0, C.data_ptr(), N, sh)
        e_ev.record(stream)
        torch.cuda.synchronize()
        times.append(s_ev.elapsed_time(e_ev))
    ms = min(times)
    return {"N": N, "ms": round(ms, 3), "quad_GFLOPs": round(2.0 * N ** 3 / (ms * 1e-3) / 1e9, 2)}


def cpu_baseline(n, m, T, threads):
    """The oracle's restatement of the reference order (level1.hpp dot with T chunks, level2.hpp
    gemv) on this host's cores: one full step of the same workload, median of 3 runs."""
    import oracle as O
    import numpy as np
    if threads:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    x = O.fill(0, 1, n)
    y = O.fill(0, 2, n)
    A = O.fill(0, 3, m * m)
    xv = O.fill(0, 4, m)
    yv0 = O.fill(0, 5, m)
    times = []
    for _ in range(3):
        yv = yv0.copy()
        t0 = time.perf_counter()
        O.dot(n, x, 1, y, 1, T)
        O.gemv(0, m, m, 1.5, A, m, xv, 1, -0.5, yv, 1)
        yv = yv0.copy()
        O.gemv(1, m, m, 1.5, A, m, xv, 1, -0.5, yv, 1)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    step_bytes = 32.0 * n + 2 * 16.0 * (m * m + m + 2 * m)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    return {"value": round(step_bytes / t / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": "one full step (qdot 2^24 + qgemv 8192^2 RowMajor + ColMajor), reference order, "
                      "oracle soft-binary128 FMA, OpenMP, median of 3",
            "seconds_per_step": round(t, 3)}


if __name__ == "__main__":
    main()
