#!/usr/bin/env python3
"""Headline benchmark: SGEMM 8192^3 (FP32, general form C = 1.5*A*B - 0.5*C,
row-major, no transpose, inputs U[-1,1] generated on the device) through the
library's C ABI on one MI355X, timed with HIP events on the library stream.

Prints ONE JSON line.  `value` is TFLOP/s counted as 2*M*N*K per step.
Also reported: the roofline fraction of the dominant kernel, a CPU baseline
(the reference buffer-kernel arithmetic ported to the host cores, on a bounded
row sample of the same product), and secondary hot-path figures (SGEMV,
SAXPY, DGEMM) in `secondary`.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3   # TFLOP/s, MI355X dense FP32 (vector or matrix)
PEAK_FP64 = 78.6    # TFLOP/s
PEAK_HBM = 8000.0   # GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=8192)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-secondary", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    return p.parse_args()


def timed(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(steps):
        fn()
    ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    return ms / steps, wall


def cpu_baseline(n, budget_s):
    """Reference buffer-kernel arithmetic (k-ascending unfused chains) on the
    host cores, on a bounded sample of output rows of the n^3 product."""
    import numpy as np

    from oracle import oracle as ora
    from tests.helpers import uniform_host

    A = uniform_host(n * n, 0x6A7A5EED, 0)
    B = uniform_host(n * n, 0x6A7A5EED, 1)
    C = uniform_host(n * n, 0x6A7A5EED, 2)
    rows = 8
    t0 = time.perf_counter()
    ora.port_matmul_rows(A, B, C, n, n, n, 0, rows, gemm=True, alpha=1.5, beta=-0.5)
    t1 = time.perf_counter() - t0
    rows2 = int(max(rows, min(n, rows * budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    ora.port_matmul_rows(A, B, C, n, n, n, 0, rows2, gemm=True, alpha=1.5, beta=-0.5)
    t2 = time.perf_counter() - t0
    tflops = 2.0 * rows2 * n * n / t2 / 1e12
    return {"value": tflops, "unit": "TFLOP/s", "cores": ora.num_threads(), "kind": "port",
            "sample": "%d of %d output rows of the %d^3 FP32 general-form product (%.1f s); reference "
                      "buffer-kernel arithmetic: k-ascending unfused multiply-add chains, OpenMP over rows"
                      % (rows2, n, n, t2)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU and the hot path is "
              "benchmarked on a single MI355X; rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from gatlas_amd import _capi
    from gatlas_amd.device import Context

    ctx = Context(0)
    n = args.n
    A, B, C = ctx.empty(n * n), ctx.empty(n * n), ctx.empty(n * n)
    ctx.fill_uniform(A, 0x6A7A5EED, 0)
    ctx.fill_uniform(B, 0x6A7A5EED, 1)
    ctx.fill_uniform(C, 0x6A7A5EED, 2)
    ctx.sync()

    def step():
        ctx.gemm(A, B, C, n, n, n, alpha=1.5, beta=-0.5)

    ms, wall = timed(ctx, step, args.steps, args.warmup)
    flops = 2.0 * n * n * n
    tflops = flops / (ms * 1e-3) / 1e12
    out = {
        "metric": "TFLOP/s (% peak) SGEMM 8192^3",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (U[-1,1] from the device generator, seed 0x6A7A5EED)",
        "config": {"workload": "SGEMM C = 1.5*A*B - 0.5*C, M=N=K=%d, FP32 row-major, no transpose" % n,
                   "M": n, "N": n, "K": n, "kernel": "MFMA FP32 (v_mfma_f32_32x32x2f32)",
                   "pct_peak_fp32": round(100.0 * tflops / PEAK_FP32, 2),
                   "wall_s_timed_region": round(wall, 4)},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_FP32, 4), "traffic": None},
    }
    for a in (A, B, C):
        a.free()

    sec = {}
    if not args.no_secondary:
        # SGEMV y = A x + y, M = N = 16384
        m = 16384
        Av, xv, yv = ctx.empty(m * m), ctx.empty(m), ctx.empty(m)
        ctx.fill_uniform(Av, 5, 0)
        ctx.fill_uniform(xv, 5, 1)
        ctx.fill_uniform(yv, 5, 2)
        for ta in (0, 1):
            ms_v, _ = timed(ctx, lambda: ctx.gemv(Av, xv, yv, m, m, transA=ta, alpha=1.0, beta=1.0), 20, 3)
            gbs = 4.0 * (m * m + m + 2 * m) / (ms_v * 1e-3) / 1e9
            sec["sgemv_16384%s" % ("_T" if ta else "")] = {"ms": round(ms_v, 4), "GB/s": round(gbs, 1),
                                                          "frac_hbm": round(gbs / PEAK_HBM, 4)}
        for a in (Av, xv, yv):
            a.free()
        # SAXPY n = 2^28
        ns = 1 << 28
        X, Y, Z = ctx.empty(ns), ctx.empty(ns), ctx.empty(ns)
        ctx.fill_uniform(X, 7, 0)
        ctx.fill_uniform(Y, 7, 1)
        ms_s, _ = timed(ctx, lambda: ctx.axpy(1.5, X, Y, Z), 20, 3)
        gbs = 12.0 * ns / (ms_s * 1e-3) / 1e9
        sec["saxpy_2^28"] = {"ms": round(ms_s, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / PEAK_HBM, 4)}
        for a in (X, Y, Z):
            a.free()
        # DGEMM 8192^3 general form
        Ad, Bd, Cd = ctx.empty(n * n, np.float64), ctx.empty(n * n, np.float64), ctx.empty(n * n, np.float64)
        ctx.fill_uniform(Ad, 3, 0)
        ctx.fill_uniform(Bd, 3, 1)
        ctx.fill_uniform(Cd, 3, 2)
        ms_d, _ = timed(ctx, lambda: ctx.gemm(Ad, Bd, Cd, n, n, n, alpha=1.5, beta=-0.5), 5, 1)
        tf = flops / (ms_d * 1e-3) / 1e12
        sec["dgemm_8192"] = {"ms": round(ms_d, 3), "TFLOP/s": round(tf, 2), "frac_fp64": round(tf / PEAK_FP64, 4)}
        for a in (Ad, Bd, Cd):
            a.free()
    out["secondary"] = sec
    ctx.close()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
