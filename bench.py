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
    p.add_argument("--no-serve", action="store_true", help="skip the serving-loop figures")
    p.add_argument("--no-grouped", action="store_true", help="skip the grouped-GEMM figures")
    p.add_argument("--serve-seconds", type=float, default=1.0, help="seconds of arrivals per serving run")
    return p.parse_args()


def measured_traffic(key):
    """HBM bytes per launch of a kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (profiles/hbm_traffic_pmc.md), scaled by the
    same counters' reading of a SAXPY of known size; None when absent."""
    import json
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            return int(json.load(f)[key]["hbm_bytes"])
    except (OSError, KeyError, ValueError):
        return None


class _Progress(dict):
    """The secondary figures; each finished entry is announced on stderr."""

    def __setitem__(self, k, v):
        super().__setitem__(k, v)
        print("[bench] %s done" % k, file=sys.stderr, flush=True)


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


def fp64_errors(got, A, B, M, N, K, rows, transA=False, transB=False):
    """Errors of sampled output rows of C = op(A)op(B) against the FP64
    product of the same FP32 operands: (componentwise max |C - C64| /
    (|op(A)||op(B)|), normwise ||C - C64||_F / ||C64||_F)."""
    import numpy as np
    a = A.reshape(K, M).T if transA else A.reshape(M, K)
    b = B.reshape(N, K).T if transB else B.reshape(K, N)
    a = np.asarray(a[rows], dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ref = a @ b
    mag = np.abs(a) @ np.abs(b)
    d = np.abs(np.asarray(got, dtype=np.float64) - ref)
    cw = float(np.max(np.where(mag > 0, d / np.where(mag > 0, mag, 1.0), d))) if d.size else 0.0
    nr = float(np.linalg.norm(ref))
    nw = float(np.linalg.norm(d) / nr) if nr > 0 else float(np.linalg.norm(d))
    return cw, nw


def _pcts(v):
    import numpy as np
    v = np.sort(np.asarray(v))
    return float(np.percentile(v, 50)), float(np.percentile(v, 99))


def grouped_bench(ctx, dtype, reps=10):
    """512 independent problems, M, N, K log-uniform in [64, 4096] (multiples
    of 16), random transposes: one tuned, graph-captured persistent launch
    against the per-problem launch loop (single-launch default tiles).  For
    FP32 operands (FP32 and split-BF16x3 forms) the componentwise error of
    4 rows of every 16th problem against FP64 is reported."""
    import numpy as np

    from gatlas_amd import _capi
    rng = np.random.default_rng(0x67525550)
    v = np.clip(np.round(np.exp(rng.uniform(np.log(64), np.log(4096), (512, 3))) / 16) * 16, 64, 4096).astype(int)
    t = rng.integers(0, 2, (512, 2))
    probs, arrays = [], []
    for i in range(512):
        M, N, K = (int(x) for x in v[i])
        ta, tb = int(t[i, 0]), int(t[i, 1])
        if dtype == _capi.GAT_BF16:
            f = ctx.empty(max(M * K, K * N))
            dA, dB = ctx.empty(M * K, np.uint16), ctx.empty(K * N, np.uint16)
            ctx.fill_uniform(f, 21, 2 * i)
            ctx.convert_bf16_n(f, dA, M * K)
            ctx.fill_uniform(f, 21, 2 * i + 1)
            ctx.convert_bf16_n(f, dB, K * N)
            f.free()
        else:
            dA, dB = ctx.empty(M * K), ctx.empty(K * N)
            ctx.fill_uniform(dA, 21, 2 * i)
            ctx.fill_uniform(dB, 21, 2 * i + 1)
        dC = ctx.empty(M * N)
        arrays += [dA, dB, dC]
        probs.append(dict(transA=ta, transB=tb, M=M, N=N, K=K, alpha=1.0, beta=0.0, A=dA, B=dB, C=dC))
    flops = sum(2.0 * p["M"] * p["N"] * p["K"] for p in probs)
    ctx.sync()

    # the per-problem loop first, with the single-launch default tile rule
    def per_problem():
        for p in probs:
            ctx.gemm(p["A"], p["B"], p["C"], p["M"], p["N"], p["K"], bool(p["transA"]), bool(p["transB"]),
                     1.0, 0.0, dtype=dtype)
    ctx.tuned_clear()
    per_problem()
    ctx.sync()
    loop = []
    for _ in range(reps):
        ctx.timer_start()
        per_problem()
        loop.append(ctx.timer_stop())
    t0 = time.perf_counter()
    nb = ctx.grouped_tune(probs, dtype)
    tune_s = time.perf_counter() - t0
    g = ctx.grouped(probs, dtype, use_graph=True)
    g.run()
    ctx.sync()
    gl = []
    for _ in range(reps):
        ctx.timer_start()
        g.run()
        gl.append(ctx.timer_stop())
    ctx.tuned_clear()
    info = g.info()
    g.close()
    err = None
    if dtype != _capi.GAT_BF16:
        err = 0.0
        for i in range(0, 512, 16):
            p = probs[i]
            M, N, K = p["M"], p["N"], p["K"]
            rows = np.unique(np.linspace(0, M - 1, 4).astype(np.int64))
            got = p["C"].download().reshape(M, N)[rows]
            cw, _ = fp64_errors(got, p["A"].download(), p["B"].download(), M, N, K, rows,
                                bool(p["transA"]), bool(p["transB"]))
            err = max(err, cw)
    for a in arrays:
        a.free()
    g50, g99 = _pcts(gl)
    l50, l99 = _pcts(loop)
    out = {"problems": 512, "gflop": round(flops / 1e9, 2), "tiles": info["tiles"], "grid": info["grid"],
           "buckets_tuned": nb, "tune_s": round(tune_s, 2),
           "grouped_TFLOP/s": round(flops / (g50 * 1e-3) / 1e12, 2), "grouped_ms_p50": round(g50, 3),
           "grouped_ms_p99": round(g99, 3),
           "loop_TFLOP/s": round(flops / (l50 * 1e-3) / 1e12, 2), "loop_ms_p50": round(l50, 3),
           "loop_ms_p99": round(l99, 3), "speedup": round(l50 / g50, 3)}
    if err is not None:
        out["max_componentwise_err"] = float("%.3e" % err)
        out["err_sample"] = "4 rows of every 16th problem vs the FP64 product of the FP32 operands"
    return out


def sgemm_1024_bench(ctx, with_cpu):
    """The reference's own runnable case: SGEMM C = A*B, M = N = K = 1024,
    FP32, no transpose, on the GPU and, timed in the same run, the reference
    buffer-kernel arithmetic ported to the host cores (whole product)."""
    import numpy as np

    n = 1024
    A, B, C = ctx.empty(n * n), ctx.empty(n * n), ctx.empty(n * n)
    ctx.fill_uniform(A, 0x6A7A5EED, 0)
    ctx.fill_uniform(B, 0x6A7A5EED, 1)
    ms, _ = timed(ctx, lambda: ctx.gemm(A, B, C, n, n, n), 200, 5)
    out = {"gpu_ms": round(ms, 4), "gpu_TFLOP/s": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 2),
           "tile(wm,wn,tm,tn,bk,blk,wk)": list(ctx.tile_for(n, n, n))}
    for a in (A, B, C):
        a.free()
    if with_cpu:
        from oracle import oracle as ora
        from tests.helpers import uniform_host
        hA, hB = uniform_host(n * n, 0x6A7A5EED, 0), uniform_host(n * n, 0x6A7A5EED, 1)
        hC = np.zeros(n * n, np.float32)
        t0 = time.perf_counter()
        ora.port_matmul_rows(hA, hB, hC, n, n, n, 0, n)
        t = time.perf_counter() - t0
        out.update({"cpu_port_s": round(t, 4), "cpu_port_TFLOP/s": round(2.0 * n ** 3 / t / 1e12, 4),
                    "cpu_threads": ora.num_threads(), "gpu_over_cpu": round(t / (ms * 1e-3), 1)})
    return out


def sweep_bench(ctx, sizes=(1024, 2048, 4096, 8192)):
    """General form C = 1.5*op(A)*op(B) - 0.5*C, all four transpose pairs,
    square sizes, FP32 and FP64 (default tile choice): TFLOP/s per case."""
    import numpy as np

    from gatlas_amd import _capi
    out = {}
    for dt, name, npdt in ((_capi.GAT_F32, "f32", np.float32), (_capi.GAT_F64, "f64", np.float64)):
        nmax = max(sizes)
        A, B, C = ctx.empty(nmax * nmax, npdt), ctx.empty(nmax * nmax, npdt), ctx.empty(nmax * nmax, npdt)
        for i, a in enumerate((A, B, C)):
            ctx.fill_uniform(a, 0x6A7A5EED, i)
        for n in sizes:
            # about 10 ms of timed work per case (at least 3 launches), after
            # 3 warm-up launches: short runs of small sizes read the clock ramp
            steps = max(3, min(400, int(1.2e12 / (2.0 * n ** 3))))
            for ta in (0, 1):
                for tb in (0, 1):
                    ms, _ = timed(ctx, lambda: ctx.gemm(A, B, C, n, n, n, bool(ta), bool(tb), 1.5, -0.5, dtype=dt),
                                  steps, 3)
                    out["%s_%d_%s%s" % (name, n, "T" if ta else "N", "T" if tb else "N")] = round(
                        2.0 * n ** 3 / (ms * 1e-3) / 1e12, 2)
        for a in (A, B, C):
            a.free()
    out["unit"] = "TFLOP/s (2n^3 per call)"
    return out


def tuner_bench(ctx, sizes=(1024, 2048, 4096)):
    """The autotuner's search over the MFMA tile table per size bucket: the
    default tile against the installed one, each timed in three alternating
    rounds after the search (median), so clock drift cannot favour either."""
    import tempfile

    out = {}
    with tempfile.TemporaryDirectory() as d:
        for n in sizes:
            A, B, C = ctx.empty(n * n), ctx.empty(n * n), ctx.empty(n * n)
            for i, a in enumerate((A, B, C)):
                ctx.fill_uniform(a, 0x6A7A5EED, i)
            # about 5 ms of launches at 100 TFLOP/s, 10..200 of them
            steps = max(10, min(200, int(5e-3 * 100e12 / (2.0 * n ** 3))))
            ctx.tuned_clear()
            dcfg = ctx.tile_for(n, n, n)
            t0 = time.perf_counter()
            cfg, _ = ctx.tune_gemm(n, n, n, journal=os.path.join(d, "tune_%d.journal" % n))
            tune_s = time.perf_counter() - t0
            ms = {"default": [], "tuned": []}
            for _ in range(3):
                for key, c in (("default", dcfg), ("tuned", cfg)):
                    ms[key].append(timed(ctx, lambda: ctx.gemm(A, B, C, n, n, n, cfg=c), steps, 5)[0])
            tf = {k: 2.0 * n ** 3 / (sorted(v)[1] * 1e-3) / 1e12 for k, v in ms.items()}
            out["sgemm_%d" % n] = {"default_TFLOP/s": round(tf["default"], 2), "tuned_TFLOP/s": round(tf["tuned"], 2),
                                   "tile_default": list(dcfg), "tile(wm,wn,tm,tn,bk,blk,wk)": list(cfg),
                                   "search_s": round(tune_s, 2)}
            for a in (A, B, C):
                a.free()
    ctx.tuned_clear()
    return out


def mixed_errors(ctx, n, nrows=32):
    """Max relative error against the FP64 product for every precision form on
    both input distributions (A, B ~ N(0,1); the same with every column of A
    and of B scaled by 10^u, u ~ U(-3,3)), on `nrows` evenly spaced output
    rows of the n^3 product C = A*B.  componentwise = max |C - C64| /
    (|A||B|), normwise = ||C - C64||_F / ||C64||_F over the sampled rows."""
    import numpy as np

    from gatlas_amd import _capi
    A, B, C = ctx.empty(n * n), ctx.empty(n * n), ctx.empty(n * n)
    rows = np.unique(np.linspace(0, n - 1, nrows).astype(np.int64))
    forms = (("f32", _capi.GAT_F32), ("bf16x3", _capi.GAT_BF16X3), ("bf16", _capi.GAT_BF16), ("fp8", _capi.GAT_FP8))
    out = {}
    for dist in ("normal", "ill_scaled_columns"):
        ctx.fill_normal(A, 0x6A7A5EED, 10)
        ctx.fill_normal(B, 0x6A7A5EED, 11)
        if dist != "normal":
            ctx.scale_columns(A, n, n, 0x6A7A5EED, 12)
            ctx.scale_columns(B, n, n, 0x6A7A5EED, 13)
        hA, hB = A.download(), B.download()
        res = {}
        for name, dt in forms:
            if dt == _capi.GAT_BF16:
                a16, b16 = ctx.empty(n * n, np.uint16), ctx.empty(n * n, np.uint16)
                ctx.convert_bf16(A, a16)
                ctx.convert_bf16(B, b16)
                ctx.gemm(a16, b16, C, n, n, n, dtype=dt)
                a16.free()
                b16.free()
            else:
                ctx.gemm(A, B, C, n, n, n, dtype=dt)
            got = C.download().reshape(n, n)[rows]
            cw, nw = fp64_errors(got, hA, hB, n, n, n, rows)
            res[name] = {"componentwise": float("%.3e" % cw), "normwise": float("%.3e" % nw)}
        out[dist] = res
    for a in (A, B, C):
        a.free()
    out["sample"] = "%d of %d output rows, C = A*B, FP64 product of the FP32 inputs" % (len(rows), n)
    return out


def serve_bench(ctx, rate, mode, seconds, gemm_dtype=0):
    """Open-loop Poisson trace: 70% GEMM (grouped-workload shapes, FP32
    operands; computed in FP32 or split-BF16x3), 30% GEMV (M, N log-uniform
    in [1024, 32768]), resident operands, 64 GB admission budget."""
    from gatlas_amd.device import Server
    srv = Server(ctx, rate=float(rate), duration=float(seconds), mode=mode, seed=rate, gemm_dtype=gemm_dtype)
    st = srv.run()
    srv.close()
    keys = ("requests", "launches", "wall_s", "tflops", "req_per_s", "p50_ms", "p99_ms", "p50_gemm_ms",
            "p99_gemm_ms", "p50_gemv_ms", "p99_gemv_ms", "mean_queue_ms", "mean_batch")
    out = {k: (round(st[k], 4) if isinstance(st[k], float) else st[k]) for k in keys}
    out["peak_footprint_GB"] = round(st["peak_footprint"] / 1e9, 2)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the hot path is benchmarked on a single MI355X "
              "(multi-GPU sharding is a design, not a measured configuration); rerun with --gpus 1",
              file=sys.stderr)
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
                   "tile(wm,wn,tm,tn,bk,blk,wk)": list(ctx.tile_for(n, n, n)),
                   "pct_peak_fp32": round(100.0 * tflops / PEAK_FP32, 2),
                   "wall_s_timed_region": round(wall, 4)},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_FP32, 4),
                     "traffic": measured_traffic("sgemm_8192_general_nn") if n == 8192 else None},
    }
    sec = _Progress()
    if not args.no_secondary:
        # mixed precision on the same operands (FP32 inputs resident in HBM)
        ms_x3, _ = timed(ctx, lambda: ctx.gemm(A, B, C, n, n, n, dtype=_capi.GAT_BF16X3), 10, 2)
        tf_x3 = flops / (ms_x3 * 1e-3) / 1e12
        sec["bf16x3_8192"] = {"ms": round(ms_x3, 4), "TFLOP/s": round(tf_x3, 2),
                              "vs_fp32_path": round(tf_x3 / tflops, 3),
                              "tile(wm,wn,tm,tn,bk,blk,wk)": list(ctx.tile_for(n, n, n, dtype=_capi.GAT_BF16X3)),
                              "note": "FP32 inputs split into bf16 hi/lo planes by a split pass (included in "
                                      "the time), 3 MFMA products per element pair; TFLOP/s counted as 2MNK"}
        a16, b16 = ctx.empty(n * n, np.uint16), ctx.empty(n * n, np.uint16)
        ctx.convert_bf16(A, a16)
        ctx.convert_bf16(B, b16)
        ms_b, _ = timed(ctx, lambda: ctx.gemm(a16, b16, C, n, n, n, dtype=_capi.GAT_BF16), 10, 2)
        tf_b = flops / (ms_b * 1e-3) / 1e12
        sec["bf16_8192"] = {"ms": round(ms_b, 4), "TFLOP/s": round(tf_b, 2), "frac_bf16_peak": round(tf_b / 2500, 4)}
        a16.free()
        b16.free()
        nb = (n // 128) * (n // 128)
        A8, B8 = ctx.empty(n * n, np.uint8), ctx.empty(n * n, np.uint8)
        sA, sB = ctx.empty(nb, np.float32), ctx.empty(nb, np.float32)
        ms_q, _ = timed(ctx, lambda: (ctx.quantize_fp8(A, n, n, n, False, A8, sA),
                                      ctx.quantize_fp8(B, n, n, n, True, B8, sB)), 3, 1)
        ms_f8, _ = timed(ctx, lambda: ctx.gemm_fp8(A8, sA, B8, sB, C, n, n, n), 10, 2)
        tf_f8 = flops / (ms_f8 * 1e-3) / 1e12
        sec["fp8_8192"] = {"ms": round(ms_f8, 4), "TFLOP/s": round(tf_f8, 2), "frac_fp8_peak": round(tf_f8 / 5000, 4),
                           "quantize_ms_both_operands": round(ms_q, 4)}
        for a in (A8, B8, sA, sB):
            a.free()
    for a in (A, B, C):
        a.free()
    if not args.no_secondary:
        sec["mixed_precision_errors_8192"] = mixed_errors(ctx, n)

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
        sec["sgemm_1024"] = sgemm_1024_bench(ctx, not args.no_cpu)
        sec["general_form_sweep"] = sweep_bench(ctx)
        sec["tuner"] = tuner_bench(ctx)
    if not args.no_grouped:
        for dt, name in ((_capi.GAT_F32, "f32"), (_capi.GAT_BF16X3, "bf16x3"), (_capi.GAT_BF16, "bf16")):
            sec["grouped512_" + name] = grouped_bench(ctx, dt)
    if not args.no_serve:
        for rate in (2000, 10000, 50000):
            for mode, mname in ((_capi.GAT_SERVE_BATCHED, "batched"), (_capi.GAT_SERVE_LOOP, "loop")):
                sec["serve_%dk_%s" % (rate // 1000, mname)] = serve_bench(ctx, rate, mode, args.serve_seconds)
            sec["serve_%dk_batched_bf16x3" % (rate // 1000)] = serve_bench(
                ctx, rate, _capi.GAT_SERVE_BATCHED, args.serve_seconds, _capi.GAT_BF16X3)
    out["secondary"] = sec
    ctx.close()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
