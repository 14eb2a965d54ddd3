"""Run-to-run determinism of every kernel family: the same inputs, launched
again and again (with other work between the launches), give the same bits.
The direct-to-LDS tiles refill a stage buffer right after the barrier that
ends its last reads, and past the last stage repeat loads into a buffer
other waves may still be reading (values never used); a missing wait or
barrier there would show as results that differ between runs.  Sizes are
chosen so interior tiles take the direct-to-LDS path for several stages
and ragged edges take the register path in the same launch."""
import ctypes

import numpy as np
import pytest

from gatlas_amd import _capi

pytestmark = pytest.mark.gpu

RUNS = 4


def _say(*a):
    print("[case]", *a, flush=True)


def _same_every_time(ctx, fn, out, clobber):
    first = None
    for i in range(RUNS):
        clobber()
        fn()
        ctx.sync()
        h = out.download()
        got = h.view(np.uint32 if h.dtype == np.float32 else np.uint64)
        if first is None:
            first = got.copy()
        else:
            assert np.array_equal(got, first), "run %d differs in %d elements" % (i, int((got != first).sum()))


def _filled(ctx, n, dt, seed, stream):
    a = ctx.empty(n, dt)
    ctx.fill_uniform(a, seed, stream)
    return a


@pytest.mark.parametrize("dtype", [_capi.GAT_F32, _capi.GAT_F64, _capi.GAT_BF16X3])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_gemm_forms_deterministic(ctx, dtype, ta, tb):
    M, N, K = 1100, 900, 1040
    _say("gemm", dtype, ta, tb, M, N, K)
    dt = np.float64 if dtype == _capi.GAT_F64 else np.float32
    A, B = _filled(ctx, M * K, dt, 41, 0), _filled(ctx, K * N, dt, 41, 1)
    C, junk = ctx.empty(M * N, dt), _filled(ctx, M * N, dt, 41, 2)
    _same_every_time(ctx, lambda: ctx.gemm(A, B, C, M, N, K, bool(ta), bool(tb), 1.5, 0.0, dtype=dtype), C,
                     lambda: C.upload(junk.download()))


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_bf16_direct_lds_deterministic(ctx, ta, tb):
    # 256 x 256 tiles staged 64 deep: 4 x 3 interior tiles and an edge strip
    M, N, K = 1100, 800, 1000
    _say("bf16", ta, tb, M, N, K)
    A32, B32 = _filled(ctx, M * K, np.float32, 43, 0), _filled(ctx, K * N, np.float32, 43, 1)
    A, B = ctx.empty(M * K, np.uint16), ctx.empty(K * N, np.uint16)
    ctx.convert_bf16(A32, A)
    ctx.convert_bf16(B32, B)
    C, junk = ctx.empty(M * N), _filled(ctx, M * N, np.float32, 43, 2)
    _same_every_time(ctx, lambda: ctx.gemm(A, B, C, M, N, K, bool(ta), bool(tb), 1.0, 0.0, dtype=_capi.GAT_BF16),
                     C, lambda: C.upload(junk.download()))


def test_fp8_deterministic(ctx):
    # every FP8 tile: 128-deep stages, interior tiles straight into LDS
    M, N, K = 1100, 800, 1024
    A32, B32 = _filled(ctx, M * K, np.float32, 47, 0), _filled(ctx, N * K, np.float32, 47, 1)
    A8, B8 = ctx.empty(M * K, np.uint8), ctx.empty(N * K, np.uint8)
    kb = K // 128
    sA, sB = ctx.empty(((M + 127) // 128) * kb), ctx.empty(((N + 127) // 128) * kb)
    ctx.quantize_fp8(A32, M, K, K, False, A8, sA)
    ctx.quantize_fp8(B32, N, K, K, False, B8, sB)
    C, junk = ctx.empty(M * N), _filled(ctx, M * N, np.float32, 47, 2)
    for cfg in ctx.tile_cfgs(_capi.GAT_FP8):
        _say("fp8", cfg)
        _capi.check(ctx.lib.gat_tuned_set(ctx.h, _capi.GAT_FP8, 0, 1, M, N, K, ctypes.byref(_capi.TileCfg(*cfg))),
                    "gat_tuned_set")
        try:
            _same_every_time(ctx, lambda: ctx.gemm_fp8(A8, sA, B8, sB, C, M, N, K), C,
                             lambda: C.upload(junk.download()))
        finally:
            ctx.tuned_clear()


@pytest.mark.parametrize("dtype", [_capi.GAT_F32, _capi.GAT_BF16X3])
def test_grouped_deterministic(ctx, dtype):
    rng = np.random.default_rng(53)
    probs = []
    for i in range(24):
        M, N, K = (int(v) for v in rng.integers(40, 700, 3))
        ta, tb = (int(v) for v in rng.integers(0, 2, 2))
        A, B = _filled(ctx, M * K, np.float32, 53, 3 * i), _filled(ctx, K * N, np.float32, 53, 3 * i + 1)
        C = ctx.empty(M * N)
        probs.append(dict(transA=ta, transB=tb, M=M, N=N, K=K, alpha=1.0, beta=0.0, A=A, B=B, C=C))
    _say("grouped", dtype)
    g = ctx.grouped(probs, dtype, use_graph=False)
    try:
        first = None
        for _ in range(RUNS):
            for p in probs:
                p["C"].upload(np.full(p["M"] * p["N"], 7.0, np.float32))
            g.run()
            ctx.sync()
            got = [p["C"].download().view(np.uint32) for p in probs]
            if first is None:
                first = got
            else:
                for a, b in zip(got, first):
                    assert np.array_equal(a, b)
    finally:
        g.close()


@pytest.mark.parametrize("ta", [False, True])
def test_gemv_deterministic(ctx, ta):
    # the transposed form sums per-block partials in a second pass
    M, N = 5000, 7000
    _say("gemv", ta, M, N)
    # op(A) is M x N (A stored N x M when transposed): x has N elements, y M
    A, x = _filled(ctx, M * N, np.float32, 59, 0), _filled(ctx, N, np.float32, 59, 1)
    y0 = _filled(ctx, M, np.float32, 59, 2)
    y = ctx.empty(M)
    _same_every_time(ctx, lambda: ctx.gemv(A, x, y, M, N, transA=ta, alpha=1.0, beta=1.0), y,
                     lambda: y.upload(y0.download()))
