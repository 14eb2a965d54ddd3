"""Runs named hot-path workloads a few times each, for profiling one kernel
at its real size (e.g. under rocprofv3 --pmc), without the benchmark's other
work:

    python3 -m gatlas_amd.kernel_runs --reps 3 sgemm8192 saxpy sgemv16384

Workloads (inputs U[-1,1] from the device generator):
  sgemm8192    C = 1.5*A*B - 0.5*C, FP32, M=N=K=8192, no transpose
  sgemm1024    C = A*B, FP32, M=N=K=1024, no transpose (default tile rule)
  dgemm8192    the same in FP64;  dgemm1024: C = A*B, FP64, 1024^3
  bf16x3_8192  split-BF16x3 (FP32 in/out), C = A*B;  bf16x3_2048 at 2048^3
  bf16_8192    BF16 inputs, FP32 out, C = A*B
  fp8_8192     resident e4m3 operands with block scales, C = A*B
  sgemv16384   y = A*x + y, FP32, 16384 x 16384 (A^T with sgemv16384t)
  saxpy        z = 1.5*x + y, FP32, n = 2^28
  grouped512   512 FP32 problems (M, N, K log-uniform in [64, 4096]) in one
               persistent grouped launch (untuned tile rule)
  grouped512x3 the same problems in split-BF16x3
  grouped512bf16 the same sizes and transposes with bf16 operands (BF16 form)

--tile W_M,W_N,T_M,T_N,BK,BLK,WK installs that tile for the GEMM workloads'
size bucket first (the tuned-table entry the autotuner would write), so one
tile of the table can be profiled at the workload's size.

Vendor reference points (PyTorch's GEMMs on the same device, i.e. the ROCm
libraries it dispatches to; timed with torch events, best of 5 rounds of 20
calls, and printed as TFLOP/s; comparison only, nothing of this library
runs, and they take a process of their own):
  vendor_sgemm8192  torch.addmm, FP32, C = 1.5*A*B - 0.5*C, 8192^3
  vendor_sgemm2048  the same at 2048^3;  vendor_sgemm1024 at 1024^3
  vendor_dgemm8192  the same in FP64;    vendor_dgemm1024 at 1024^3
  vendor_bf16_8192  torch.matmul of bf16 operands (bf16 output), 8192^3
  vendor_fp8_8192   torch._scaled_mm, e4m3 operands, unit tensor scales,
                    FP32 output, 8192^3 (vendor_fp8_8192_bf16out: bf16 out)
"""
import argparse
import sys

import numpy as np

from . import _capi
from .device import Context


def _sq(ctx, n, dt=np.float32, count=3):
    arrs = [ctx.empty(n * n, dt) for _ in range(count)]
    for i, a in enumerate(arrs):
        ctx.fill_uniform(a, 0x6A7A5EED, i)
    return arrs


# GEMM workloads: (dtype, transA, transB, M, N, K) of their tuned-table bucket
# (the FP8 product reads B as N x K, i.e. op(B) transposed)
_GEMM_SHAPES = {
    "sgemm8192": (_capi.GAT_F32, 0, 0, 8192, 8192, 8192),
    "sgemm1024": (_capi.GAT_F32, 0, 0, 1024, 1024, 1024),
    "dgemm8192": (_capi.GAT_F64, 0, 0, 8192, 8192, 8192),
    "dgemm1024": (_capi.GAT_F64, 0, 0, 1024, 1024, 1024),
    "bf16x3_8192": (_capi.GAT_BF16X3, 0, 0, 8192, 8192, 8192),
    "bf16x3_2048": (_capi.GAT_BF16X3, 0, 0, 2048, 2048, 2048),
    "bf16_8192": (_capi.GAT_BF16, 0, 0, 8192, 8192, 8192),
    "fp8_8192": (_capi.GAT_FP8, 0, 1, 8192, 8192, 8192),
}


def install_tile(ctx, name, tile):
    import ctypes
    if name not in _GEMM_SHAPES:
        raise SystemExit("--tile applies to the GEMM workloads only, not %s" % name)
    dt, ta, tb, M, N, K = _GEMM_SHAPES[name]
    _capi.check(ctx.lib.gat_tuned_set(ctx.h, dt, ta, tb, M, N, K, ctypes.byref(_capi.TileCfg(*tile))),
                "gat_tuned_set")


def vendor(name):
    """TFLOP/s of PyTorch's own GEMM for one vendor_* workload."""
    import torch
    dev = "cuda:0"
    g = torch.Generator(device=dev).manual_seed(0x6A7A5EED)

    def uni(n, dt):
        return (torch.rand(n, n, device=dev, generator=g, dtype=torch.float32) * 2 - 1).to(dt)

    import re
    kind = name.split("_")[1]
    n = int(re.search(r"\d{3,}", name).group(0))  # the size: the first run of 3+ digits
    if kind.startswith("sgemm") or kind.startswith("dgemm"):
        dt = torch.float32 if kind[0] == "s" else torch.float64
        A, B, C = uni(n, dt), uni(n, dt), uni(n, dt)
        fn = lambda: torch.addmm(C, A, B, beta=-0.5, alpha=1.5)  # noqa: E731
    elif kind == "bf16":
        A, B = uni(n, torch.bfloat16), uni(n, torch.bfloat16)
        fn = lambda: A @ B  # noqa: E731
    elif kind == "fp8":
        A8 = uni(n, torch.float32).to(torch.float8_e4m3fn)
        B8 = uni(n, torch.float32).to(torch.float8_e4m3fn).t().contiguous().t()  # column-major operand
        one = torch.tensor(1.0, device=dev)
        out = torch.bfloat16 if name.endswith("bf16out") else torch.float32
        fn = lambda: torch._scaled_mm(A8, B8, scale_a=one, scale_b=one, out_dtype=out)  # noqa: E731
    else:
        raise SystemExit("unknown workload %s" % name)
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(5):
        s.record()
        for _ in range(20):
            fn()
        e.record()
        torch.cuda.synchronize()
        best = min(best, s.elapsed_time(e) / 20)
    tf = 2.0 * n ** 3 / (best * 1e-3) / 1e12
    print("[kernel_runs] %s %.4f ms %.1f TFLOP/s" % (name, best, tf), flush=True)
    return tf


def run(ctx, name, reps):
    n = 8192
    if name in ("sgemm1024", "dgemm1024"):
        m = 1024
        A, B, C = _sq(ctx, m, np.float32 if name[0] == "s" else np.float64)
        for _ in range(reps):
            ctx.gemm(A, B, C, m, m, m)
    elif name == "bf16x3_2048":
        m = 2048
        A, B, C = _sq(ctx, m)
        for _ in range(reps):
            ctx.gemm(A, B, C, m, m, m, dtype=_capi.GAT_BF16X3)
    elif name in ("sgemm8192", "dgemm8192"):
        A, B, C = _sq(ctx, n, np.float32 if name[0] == "s" else np.float64)
        for _ in range(reps):
            ctx.gemm(A, B, C, n, n, n, alpha=1.5, beta=-0.5)
    elif name in ("bf16x3_8192", "bf16_8192"):
        A, B, C = _sq(ctx, n)
        if name.startswith("bf16x3"):
            for _ in range(reps):
                ctx.gemm(A, B, C, n, n, n, dtype=_capi.GAT_BF16X3)
        else:
            a16, b16 = ctx.empty(n * n, np.uint16), ctx.empty(n * n, np.uint16)
            ctx.convert_bf16(A, a16)
            ctx.convert_bf16(B, b16)
            for _ in range(reps):
                ctx.gemm(a16, b16, C, n, n, n, dtype=_capi.GAT_BF16)
    elif name == "fp8_8192":
        A, B, C = _sq(ctx, n)
        nb = (n // 128) ** 2
        A8, B8 = ctx.empty(n * n, np.uint8), ctx.empty(n * n, np.uint8)
        sA, sB = ctx.empty(nb), ctx.empty(nb)
        ctx.quantize_fp8(A, n, n, n, False, A8, sA)
        ctx.quantize_fp8(B, n, n, n, True, B8, sB)
        for _ in range(reps):
            ctx.gemm_fp8(A8, sA, B8, sB, C, n, n, n)
    elif name in ("sgemv16384", "sgemv16384t"):
        m = 16384
        A = ctx.empty(m * m)
        x, y = ctx.empty(m), ctx.empty(m)
        for i, a in enumerate((A, x, y)):
            ctx.fill_uniform(a, 5, i)
        for _ in range(reps):
            ctx.gemv(A, x, y, m, m, transA=name.endswith("t"), alpha=1.0, beta=1.0)
    elif name in ("grouped512", "grouped512x3", "grouped512bf16"):
        # 512 problems, M, N, K log-uniform in [64, 4096] (multiples of 16),
        # random transposes, FP32; one persistent launch per repetition
        rng = np.random.default_rng(0x67525550)
        v = np.clip(np.round(np.exp(rng.uniform(np.log(64), np.log(4096), (512, 3))) / 16) * 16, 64,
                    4096).astype(int)
        tr = rng.integers(0, 2, (512, 2))
        bf = name.endswith("bf16")
        probs = []
        for i in range(512):
            M, N, K = (int(x) for x in v[i])
            dA, dB, dC = ctx.empty(M * K), ctx.empty(K * N), ctx.empty(M * N)
            ctx.fill_uniform(dA, 21, 2 * i)
            ctx.fill_uniform(dB, 21, 2 * i + 1)
            if bf:
                hA, hB = ctx.empty(M * K, np.uint16), ctx.empty(K * N, np.uint16)
                ctx.convert_bf16_n(dA, hA, M * K)
                ctx.convert_bf16_n(dB, hB, K * N)
                dA.free()
                dB.free()
                dA, dB = hA, hB
            probs.append(dict(transA=int(tr[i, 0]), transB=int(tr[i, 1]), M=M, N=N, K=K, A=dA, B=dB, C=dC))
        form = _capi.GAT_BF16X3 if name.endswith("x3") else _capi.GAT_BF16 if bf else _capi.GAT_F32
        g = ctx.grouped(probs, form, use_graph=False)
        for _ in range(reps):
            g.run()
        ctx.sync()
        g.close()
    elif name == "saxpy":
        ns = 1 << 28
        X, Y, Z = ctx.empty(ns), ctx.empty(ns), ctx.empty(ns)
        ctx.fill_uniform(X, 7, 0)
        ctx.fill_uniform(Y, 7, 1)
        for _ in range(reps):
            ctx.axpy(1.5, X, Y, Z)
    else:
        raise SystemExit("unknown workload %s" % name)
    ctx.sync()
    print("[kernel_runs] %s x%d done" % (name, reps), flush=True)


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--reps", type=int, default=3)
    p.add_argument("--tile", default=None, help="W_M,W_N,T_M,T_N,BK,BLK,WK for the GEMM workloads")
    p.add_argument("workloads", nargs="*", help="workload names")
    a = p.parse_args(argv)
    tile = None if a.tile is None else tuple(int(v) for v in a.tile.split(","))
    if tile is not None and len(tile) != 7:
        raise SystemExit("--tile takes 7 integers")
    vend = [w.startswith("vendor_") for w in a.workloads]
    if any(vend):
        # PyTorch brings a HIP runtime of its own, which cannot open the
        # device in a process where this library's runtime already has
        if not all(vend):
            raise SystemExit("vendor_* workloads run in a process of their own")
        for w in a.workloads:
            vendor(w)
        return 0
    ctx = Context(0)
    for w in a.workloads:
        if tile is not None:
            install_tile(ctx, w, tile)
        run(ctx, w, a.reps)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
