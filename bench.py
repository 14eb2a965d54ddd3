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


MEDIAN_NOTE = "median of 5 rounds of 10 calls, after 3 untimed calls"


def timed_median(ctx, fn, steps, rounds=5, warmup=3):
    """Median over `rounds` of the per-call device time of `steps` calls in a
    row (events on the library stream), after `warmup` untimed calls: the
    secondary GEMMs' figure, steadier than one round against the chip's
    clock changes under dense matrix work."""
    for _ in range(warmup):
        fn()
    ctx.sync()
    per = []
    for _ in range(rounds):
        ctx.timer_start()
        for _ in range(steps):
            fn()
        per.append(ctx.timer_stop() / steps)
    ctx.sync()
    return sorted(per)[len(per) // 2]


def _host_info():
    """CPU model, host threads and what clinfo reports about OpenCL devices
    (the reference's intended CPU baseline needs a CPU OpenCL device)."""
    import shutil
    import subprocess
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    probe = "clinfo not installed"
    if shutil.which("clinfo"):
        try:
            txt = subprocess.run(["clinfo"], capture_output=True, text=True, timeout=60).stdout
            types = [l.split("Device Type", 1)[1].strip() for l in txt.splitlines() if "Device Type" in l]
            plats = [l for l in txt.splitlines() if l.strip().startswith("Number of platforms")]
            probe = "%s; device types: %s; CPU OpenCL device: %s" % (
                plats[0].strip() if plats else "no platform line", ", ".join(types) or "none",
                "yes" if any("CPU" in t for t in types) else "no")
        except (OSError, subprocess.SubprocessError) as e:
            probe = "clinfo failed: %s" % e
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "clinfo": probe}


def _uniform_into(out, seed, stream, chunk=1 << 24):
    """The device generator's values (gat_fill_uniform) on the host, chunk by chunk."""
    import numpy as np

    from tests.helpers import uniform_host
    for i0 in range(0, out.size, chunk):
        i1 = min(out.size, i0 + chunk)
        out[i0:i1] = uniform_host(0, seed, stream, idx=np.arange(i0, i1, dtype=np.uint64))
    return out


def _trials(fn, trials):
    """The reference's timing: one dummy run (GatlasAppUtil.cpp:203-208),
    then `trials` timed runs (wall clock around the compute only)."""
    fn()
    ts = []
    for _ in range(trials):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return ts


def _rates(work, ts, scale):
    """Median and mean of the per-trial rates (GatlasAppUtil.cpp:252-264
    averages per-trial GFLOPS)."""
    import numpy as np
    r = [work / t / scale for t in ts]
    return float(np.median(r)), float(np.mean(r))


def cpu_baseline(n, budget_s, gpu, trials=5):
    """The reference buffer-kernel arithmetic (k-ascending unfused multiply-add
    chains per output, KernelMatmulBuffer.hpp:259-337; the GEMV and SAXPY
    kernels' chains likewise), ported to C with OpenMP over output rows, on
    the host cores of the GPU box, on the GPU run's own seeded inputs.  No CPU
    OpenCL runtime exists in this image (clinfo is recorded), so the port
    stands in for the reference's kernels on a CPU OpenCL device.  Each leg:
    one dummy run, then `trials` timed runs; median and mean of the per-trial
    rates.  The headline value is SGEMM n^3 in the general form on a bounded
    sample of output rows (about budget_s of CPU work in all)."""
    import numpy as np

    from oracle import oracle as ora
    info = _host_info()
    legs = {}

    inputs = {}

    def gemm_leg(m, ta, tb, general, rows=None):
        if m not in inputs:
            inputs.clear()
            inputs[m] = [_uniform_into(np.empty(m * m, np.float32), 0x6A7A5EED, s) for s in range(3)]
        A, B, C = inputs[m]
        r = m if rows is None else rows
        kw = dict(gemm=general, alpha=1.5, beta=-0.5) if general else {}
        ts = _trials(lambda: ora.port_matmul_rows(A, B, C, m, m, m, 0, r, transA=ta, transB=tb, **kw), trials)
        med, mean = _rates(2.0 * r * m * m, ts, 1e12)
        return {"TFLOP/s_median": round(med, 5), "TFLOP/s_mean": round(mean, 5), "trial_s": round(float(np.median(ts)), 4),
                "rows": r}

    # headline: rows of the n^3 general-form product, sized from an 8-row run
    A = _uniform_into(np.empty(n * n, np.float32), 0x6A7A5EED, 0)
    B = _uniform_into(np.empty(n * n, np.float32), 0x6A7A5EED, 1)
    C = _uniform_into(np.empty(n * n, np.float32), 0x6A7A5EED, 2)
    t0 = time.perf_counter()
    ora.port_matmul_rows(A, B, C, n, n, n, 0, 8, gemm=True, alpha=1.5, beta=-0.5)
    per_row = (time.perf_counter() - t0) / 8
    rows = int(max(8, min(n, budget_s / (trials + 1) / max(per_row, 1e-6))))
    ts = _trials(lambda: ora.port_matmul_rows(A, B, C, n, n, n, 0, rows, gemm=True, alpha=1.5, beta=-0.5), trials)
    med, mean = _rates(2.0 * rows * n * n, ts, 1e12)
    del A, B, C
    # the reference's own runnable case and the 1024/2048 sweep points
    legs["sgemm_1024_pure"] = gemm_leg(1024, 0, 0, False)
    for m in (1024, 2048):
        for ta in (0, 1):
            for tb in (0, 1):
                legs["general_%d_%s%s" % (m, "T" if ta else "N", "T" if tb else "N")] = gemm_leg(m, ta, tb, True)
    inputs.clear()
    # SGEMV y = A x + y, 16384^2 (reference chain: one row per work-item)
    m = 16384
    Av = _uniform_into(np.empty(m * m, np.float32), 5, 0)
    xv = _uniform_into(np.empty(m, np.float32), 5, 1)
    yv = _uniform_into(np.empty(m, np.float32), 5, 2)
    ts = _trials(lambda: ora.port_matvec_rows(Av, xv, yv, m, 0, m, gemv=True, alpha=1.0, beta=1.0), trials)
    med_v, mean_v = _rates(4.0 * (m * m + 3 * m), ts, 1e9)
    legs["sgemv_16384"] = {"GB/s_median": round(med_v, 2), "GB/s_mean": round(mean_v, 2),
                           "trial_s": round(float(np.median(ts)), 4)}
    del Av
    # SAXPY z = 1.5 x + y, n = 2^28 (separate z, as the reference writes it)
    ns = 1 << 28
    X = _uniform_into(np.empty(ns, np.float32), 7, 0)
    Y = _uniform_into(np.empty(ns, np.float32), 7, 1)
    Z = np.empty(ns, np.float32)
    ts = _trials(lambda: ora.port_saxpy(1.5, X, Y, Z), trials)
    med_s, mean_s = _rates(12.0 * ns, ts, 1e9)
    legs["saxpy_2^28"] = {"GB/s_median": round(med_s, 2), "GB/s_mean": round(mean_s, 2),
                          "trial_s": round(float(np.median(ts)), 4)}
    del X, Y, Z
    # GPU over CPU on the same workloads (GPU figures from this run)
    ratio = {}
    if gpu.get("sgemm_1024"):
        ratio["sgemm_1024_pure"] = round(gpu["sgemm_1024"] / legs["sgemm_1024_pure"]["TFLOP/s_median"], 1)
    if gpu.get("sgemv_16384"):
        ratio["sgemv_16384"] = round(gpu["sgemv_16384"] / legs["sgemv_16384"]["GB/s_median"], 1)
    if gpu.get("saxpy"):
        ratio["saxpy_2^28"] = round(gpu["saxpy"] / legs["saxpy_2^28"]["GB/s_median"], 1)
    for k, v in gpu.get("sweep", {}).items():
        leg = "general_" + k[4:]
        if leg in legs:
            ratio[leg] = round(v / legs[leg]["TFLOP/s_median"], 1)
    out = {"value": round(med, 6), "unit": "TFLOP/s", "cores": ora.num_threads(), "kind": "port",
           "sample": "%d of %d output rows of the %d^3 FP32 general-form product per trial; one dummy run, "
                     "then %d trials (median; mean %.6f TFLOP/s); reference buffer-kernel arithmetic: "
                     "k-ascending unfused multiply-add chains, OpenMP over rows" % (rows, n, n, trials, mean),
           "trials": trials, "TFLOP/s_mean": round(mean, 6)}
    out.update(info)
    out["legs"] = legs
    out["gpu_over_cpu"] = ratio
    return out


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
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        jnl = os.path.join(d, "grouped.jnl")
        t0 = time.perf_counter()
        nb = ctx.grouped_tune(probs, dtype, journal=jnl)
        tune_s = time.perf_counter() - t0
        with open(jnl) as f:
            recs = [ln.split() for ln in f if ln.strip()]
    trials = sum(1 for r in recs if int(r[1]) >= 0)
    g = ctx.grouped(probs, dtype, use_graph=True)
    g.run()
    ctx.sync()
    shapes = g.shapes()
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
           "buckets_tuned": nb, "tune_s": round(tune_s, 2), "trials_journaled": trials,
           "interior_shapes": {name: shapes.count(i) for i, name in enumerate(("128x128", "64x64", "128x64",
                                                                                 "64x128"))},
           "grouped_TFLOP/s": round(flops / (g50 * 1e-3) / 1e12, 2), "grouped_ms_p50": round(g50, 3),
           "grouped_ms_p99": round(g99, 3),
           "loop_TFLOP/s": round(flops / (l50 * 1e-3) / 1e12, 2), "loop_ms_p50": round(l50, 3),
           "loop_ms_p99": round(l99, 3), "speedup": round(l50 / g50, 3)}
    if err is not None:
        out["max_componentwise_err"] = float("%.3e" % err)
        out["err_sample"] = "4 rows of every 16th problem vs the FP64 product of the FP32 operands"
    return out


def sgemm_1024_bench(ctx):
    """The reference's own runnable case: SGEMM C = A*B, M = N = K = 1024,
    FP32, no transpose (its CPU counterpart is a cpu_baseline leg)."""
    n = 1024
    A, B, C = ctx.empty(n * n), ctx.empty(n * n), ctx.empty(n * n)
    ctx.fill_uniform(A, 0x6A7A5EED, 0)
    ctx.fill_uniform(B, 0x6A7A5EED, 1)
    ms, _ = timed(ctx, lambda: ctx.gemm(A, B, C, n, n, n), 200, 5)
    out = {"gpu_ms": round(ms, 4), "gpu_TFLOP/s": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 2),
           "tile(wm,wn,tm,tn,bk,blk,wk)": list(ctx.tile_for(n, n, n))}
    for a in (A, B, C):
        a.free()
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


def _journal_records(path):
    """Timed trials a journal holds (records with a value >= 0), and keys."""
    n, keys = 0, set()
    with open(path) as f:
        for line in f:
            k, _, v = line.rstrip("\n").rpartition("\t")
            if v.lstrip("-").isdigit() and int(v) >= 0:
                n += 1
                keys.add(k)
    return n, len(keys)


def tuner_bench(ctx, sizes=(1024, 2048, 4096, 8192)):
    """The autotuning searches at the workloads' sizes, journaled:
      * gat_tune_gemm (times every instantiated MFMA tile of the form, then
        confirms the winner as the EM fixed point) for FP32 and FP64, all four
        transpose pairs, general form, square sizes; the default tile and the
        installed one are then timed in three alternating rounds (median);
      * the reference's EM search (benchMatvec.cpp:521-612,
        benchSaxpy.cpp:498-584) over the GEMV y = A*x + y at 16384^2 (both
        storage orders) and SAXPY at n = 2^28, for the bandwidth kernels and
        the reference-geometry buffer kernels, through the parameter
        interface.
    Per case: search seconds, candidates journaled, best parameters."""
    import tempfile

    import numpy as np

    from gatlas_amd import _capi
    from gatlas_amd.interface import BUFFER, MATVEC, NATIVE, SAXPY, Kernel
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for dt, name, npdt in ((_capi.GAT_F32, "f32", np.float32), (_capi.GAT_F64, "f64", np.float64)):
            nmax = max(sizes)
            A, B, C = (ctx.empty(nmax * nmax, npdt) for _ in range(3))
            for i, a in enumerate((A, B, C)):
                ctx.fill_uniform(a, 0x6A7A5EED, i)
            for n in sizes:
                steps = max(3, min(200, int(5e-3 * 100e12 / (2.0 * n ** 3))))
                for ta in (0, 1):
                    for tb in (0, 1):
                        ctx.tuned_clear()
                        dcfg = ctx.tile_for(n, n, n, bool(ta), bool(tb), dt)
                        jp = os.path.join(d, "tune_%s_%d_%d%d.journal" % (name, n, ta, tb))
                        t0 = time.perf_counter()
                        cfg, _ = ctx.tune_gemm(n, n, n, bool(ta), bool(tb), dtype=dt, journal=jp)
                        tune_s = time.perf_counter() - t0
                        ms = {"default": [], "tuned": []}
                        for _ in range(3):
                            for key, c in (("default", dcfg), ("tuned", cfg)):
                                ms[key].append(timed(ctx, lambda: ctx.gemm(A, B, C, n, n, n, bool(ta), bool(tb), 1.5,
                                                                           -0.5, dtype=dt, cfg=c), steps, 2)[0])
                        tf = {k: 2.0 * n ** 3 / (sorted(v)[1] * 1e-3) / 1e12 for k, v in ms.items()}
                        trials, keys = _journal_records(jp)
                        out["%s_%d_%s%s" % (name, n, "T" if ta else "N", "T" if tb else "N")] = {
                            "default_TFLOP/s": round(tf["default"], 2), "tuned_TFLOP/s": round(tf["tuned"], 2),
                            "tile_default": list(dcfg), "tile_tuned": list(cfg), "search_s": round(tune_s, 2),
                            "candidates_journaled": keys, "trials_journaled": trials}
            for a in (A, B, C):
                a.free()
        ctx.tuned_clear()
        m, ns = 16384, 1 << 28
        searches = [("sgemv_16384_native", MATVEC, NATIVE, 4, dict(M=m, N=m, generalized=True)),
                    ("sgemv_16384_T_native", MATVEC, NATIVE, 4, dict(M=m, N=m, transA=True, generalized=True)),
                    ("sgemv_16384_buffer_float4", MATVEC, BUFFER, 4, dict(M=m, N=m, generalized=True)),
                    ("saxpy_2^28_native", SAXPY, NATIVE, 4, dict(M=ns // 4, N=4)),
                    ("saxpy_2^28_buffer_float4", SAXPY, BUFFER, 4, dict(M=ns // 4, N=4))]
        for key, fam, var, vl, kw in searches:
            k = Kernel(fam, _capi.GAT_F32, vl, var, ctx=ctx)
            jp = os.path.join(d, key + ".journal")
            lim = dict(max_block_height=64, max_group_size=256) if fam == MATVEC else \
                dict(max_block_height=16, max_group_size=256)
            t0 = time.perf_counter()
            best = k.em_search(jp, vl=vl, max_workgroup=ctx.max_workgroup(), **lim, **kw)
            search_s = time.perf_counter() - t0
            trials, keys = _journal_records(jp)
            res = {"search_s": round(search_s, 2), "candidates_journaled": keys, "trials_journaled": trials,
                   "best_params": best}
            if best:
                from gatlas_amd.interface import Journal
                j = Journal(jp)
                j.loadMemo()
                us = j.memoTime(k, best, 0)
                if us and us > 0:
                    nbytes = 4.0 * (m * m + 3 * m) if fam == MATVEC else 12.0 * ns
                    res["best_us"] = us
                    res["best_GB/s"] = round(nbytes / us / 1e3, 1)
            out[key] = res
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


def serve_bench(ctx, rate, mode, seconds, gemm_dtype=0, keep=20):
    """Open-loop Poisson trace: 70% GEMM (grouped-workload shapes, FP32
    operands; computed in FP32 or split-BF16x3), 30% GEMV (M, N log-uniform
    in [1024, 32768]), resident operands, 64 GB admission budget (operands
    count as views into the resident pool, results as arena allocations).
    About `keep` results spread over the trace are kept: each is recomputed
    alone (bit-identical required) and a row sample of each is compared with
    the FP64 product (numpy) of its operands."""
    import numpy as np

    from gatlas_amd.device import Server
    nreq = rate * seconds
    srv = Server(ctx, rate=float(rate), duration=float(seconds), mode=mode, seed=rate, gemm_dtype=gemm_dtype,
                 keep=max(1, int(nreq // keep)), keep_max=keep)
    st = srv.run()
    mism = srv.check_alone()
    worst = 0.0
    for r in srv.kept():
        M, N, K = r["M"], r["N"], r["K"]
        rows = np.unique(np.linspace(0, M - 1, min(M, 4)).astype(np.int64))
        if r["kind"] == 0:
            A, B = srv.read(r["A"], M * K), srv.read(r["B"], K * N)
            got = srv.read(r["C"], M * N).reshape(M, N)[rows]
            cw, _ = fp64_errors(got, A, B, M, N, K, rows, bool(r["transA"]), bool(r["transB"]))
        else:
            A = np.concatenate([srv.read(r["A"] + 4 * int(i) * r["lda"], N) for i in rows])
            x = srv.read(r["B"], N)
            got = srv.read(r["C"], M)[rows]
            cw, _ = fp64_errors(got[:, None], A, x, len(rows), 1, N, np.arange(len(rows)))
        worst = max(worst, cw)
    srv.close()
    keys = ("requests", "launches", "wall_s", "tflops", "offered_tflops", "req_per_s", "p50_ms", "p99_ms",
            "p50_gemm_ms", "p99_gemm_ms", "p50_gemv_ms", "p99_gemv_ms", "mean_queue_ms", "mean_batch",
            "device_busy", "blocked_ms")
    out = {k: (round(st[k], 4) if isinstance(st[k], float) else st[k]) for k in keys}
    out["peak_footprint_GB"] = round(st["peak_footprint"] / 1e9, 2)
    out["peak_output_GB"] = round(st["peak_output_bytes"] / 1e9, 3)
    out["kept_checked"] = st["kept"]
    out["kept_alone_mismatches"] = mism
    out["kept_max_componentwise_err"] = float("%.3e" % worst)
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
        ms_x3 = timed_median(ctx, lambda: ctx.gemm(A, B, C, n, n, n, dtype=_capi.GAT_BF16X3), 10)
        tf_x3 = flops / (ms_x3 * 1e-3) / 1e12
        sec["bf16x3_8192"] = {"ms": round(ms_x3, 4), "TFLOP/s": round(tf_x3, 2),
                              "vs_fp32_path": round(tf_x3 / tflops, 3), "timing": MEDIAN_NOTE,
                              "tile(wm,wn,tm,tn,bk,blk,wk)": list(ctx.tile_for(n, n, n, dtype=_capi.GAT_BF16X3)),
                              "note": "FP32 inputs split into bf16 hi/lo planes by a split pass (included in "
                                      "the time), 3 MFMA products per element pair; TFLOP/s counted as 2MNK"}
        a16, b16 = ctx.empty(n * n, np.uint16), ctx.empty(n * n, np.uint16)
        ctx.convert_bf16(A, a16)
        ctx.convert_bf16(B, b16)
        ms_b = timed_median(ctx, lambda: ctx.gemm(a16, b16, C, n, n, n, dtype=_capi.GAT_BF16), 10)
        tf_b = flops / (ms_b * 1e-3) / 1e12
        sec["bf16_8192"] = {"ms": round(ms_b, 4), "TFLOP/s": round(tf_b, 2), "frac_bf16_peak": round(tf_b / 2500, 4),
                            "timing": MEDIAN_NOTE}
        a16.free()
        b16.free()
        nb = (n // 128) * (n // 128)
        A8, B8 = ctx.empty(n * n, np.uint8), ctx.empty(n * n, np.uint8)
        sA, sB = ctx.empty(nb, np.float32), ctx.empty(nb, np.float32)
        ms_q, _ = timed(ctx, lambda: (ctx.quantize_fp8(A, n, n, n, False, A8, sA),
                                      ctx.quantize_fp8(B, n, n, n, True, B8, sB)), 3, 1)
        ms_f8 = timed_median(ctx, lambda: ctx.gemm_fp8(A8, sA, B8, sB, C, n, n, n), 10)
        tf_f8 = flops / (ms_f8 * 1e-3) / 1e12
        sec["fp8_8192"] = {"ms": round(ms_f8, 4), "TFLOP/s": round(tf_f8, 2), "frac_fp8_peak": round(tf_f8 / 5000, 4),
                           "quantize_ms_both_operands": round(ms_q, 4), "timing": MEDIAN_NOTE}
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
        sec["sgemm_1024"] = sgemm_1024_bench(ctx)
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
        gpu = {"sgemm_1024": sec.get("sgemm_1024", {}).get("gpu_TFLOP/s"),
               "sgemv_16384": sec.get("sgemv_16384", {}).get("GB/s"),
               "saxpy": sec.get("saxpy_2^28", {}).get("GB/s"),
               "sweep": {k: v for k, v in sec.get("general_form_sweep", {}).items()
                         if k.startswith("f32_1024_") or k.startswith("f32_2048_")}}
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds, gpu)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
