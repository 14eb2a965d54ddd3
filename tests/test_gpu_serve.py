"""Grouped GEMV and the serving loop on the GPU.

  * grouped GEMV: every problem bit-identical to its single gat_gemv launch
    and within 1e-5 (componentwise) of the FP64 oracle;
  * serving (small trace, GEMM requests in FP32 and in split-BF16x3): every
    request finishes, statistics are consistent, each kept result is
    bit-identical to the same request run alone, the batched and the
    per-request-loop modes give identical bits for the same trace, and kept
    results match the FP64 oracle within 1e-5;
  * a budget just above the largest request still drains the trace.
"""
import numpy as np
import pytest

from gatlas_amd import _capi
from gatlas_amd.device import Server
from oracle import oracle as ora
from tests.helpers import uniform_host

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def test_gemv_grouped_bit_identical(ctx):
    rng = np.random.default_rng(31)
    dims = [(int(m), int(n)) for m, n in rng.integers(1, 3000, (14, 2))]
    dims += [(0, 10), (7, 0), (4096, 4096), (33, 1027), (1, 1)]
    _say("gemv grouped", len(dims))
    probs, host = [], []
    for i, (M, N) in enumerate(dims):
        A = uniform_host(M * N, 41, 3 * i)
        x = uniform_host(N, 41, 3 * i + 1)
        y0 = uniform_host(M, 41, 3 * i + 2)
        probs.append(dict(M=M, N=N, alpha=1.5, beta=-0.5, A=ctx.to_device(A), x=ctx.to_device(x),
                          y=ctx.to_device(y0)))
        host.append((A, x, y0))
    ctx.gemv_grouped(probs)
    ctx.sync()
    for (M, N), (A, x, y0), p in zip(dims, host, probs):
        got = p["y"].download()
        p["y"].upload(y0)
        ctx.gemv(p["A"], p["x"], p["y"], M, N, False, 1.5, -0.5)
        ctx.sync()
        assert np.array_equal(got.view(np.uint32), p["y"].download().view(np.uint32)), (M, N)
        if M and N:
            ref, mag = ora.gemv64(A, x, M, N, False, 1.5, -0.5, y0)
            cw, _, _ = ora.error_metrics(got, ref, mag)
            assert cw <= TOL, (M, N, cw)


SMALL = dict(rate=3000.0, duration=0.15, gemm_min=64, gemm_max=512, gemv_min=256, gemv_max=4096,
             hbm_budget=1 << 28, keep=7, keep_max=48, seed=5, streams=3, max_batch=64)


def _kept_results(srv):
    out = []
    for r in srv.kept():
        n = r["M"] * r["N"] if r["kind"] == 0 else r["M"]
        out.append((r, srv.read(r["C"], n)))
    return out


def _oracle_check(srv, kept):
    for r, got in kept[:24]:
        M, N, K = r["M"], r["N"], r["K"]
        if r["kind"] == 0:
            ta, tb = bool(r["transA"]), bool(r["transB"])
            A = srv.read(r["A"], M * K)  # lda == M (ta) or K: dense
            B = srv.read(r["B"], K * N)
            rows = np.unique(np.linspace(0, M - 1, min(M, 8)).astype(np.int64))
            ref, mag = ora.gemm64(A, B, M, N, K, ta, tb, 1.0, 0.0, None, rows=rows)
            cw, _, _ = ora.error_metrics(np.ascontiguousarray(got.reshape(M, N)[rows]), ref, mag)
        else:
            lda = r["lda"]
            Ap = srv.read(r["A"], (M - 1) * lda + N).astype(np.float32)
            A = np.lib.stride_tricks.as_strided(Ap, (M, N), (4 * lda, 4)).copy()
            x = srv.read(r["B"], N)
            ref, mag = ora.gemv64(A.ravel(), x, M, N, False, 1.0, 0.0, None)
            cw, _, _ = ora.error_metrics(got, ref, mag)
        assert cw <= TOL, (r, cw)


@pytest.mark.parametrize("gemm_dtype", [_capi.GAT_F32, _capi.GAT_BF16X3], ids=["f32", "bf16x3"])
def test_serve_batched_vs_loop(ctx, gemm_dtype):
    cfg = dict(SMALL, gemm_dtype=gemm_dtype)
    _say("serve small batched", gemm_dtype)
    srv = Server(ctx, mode=_capi.GAT_SERVE_BATCHED, **cfg)
    st = srv.run()
    print("[serve batched]", {k: round(v, 4) if isinstance(v, float) else v for k, v in st.items()}, flush=True)
    assert st["requests"] == st["gemm_requests"] + st["gemv_requests"] > 100
    assert st["launches"] <= st["requests"]
    assert 0 <= st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["peak_footprint"] <= SMALL["hbm_budget"]
    assert st["kept"] > 10
    assert srv.check_alone() == 0
    kb = _kept_results(srv)
    _oracle_check(srv, kb)
    srv.close()

    _say("serve small loop", gemm_dtype)
    srv2 = Server(ctx, mode=_capi.GAT_SERVE_LOOP, **cfg)
    st2 = srv2.run()
    print("[serve loop]", {k: round(v, 4) if isinstance(v, float) else v for k, v in st2.items()}, flush=True)
    assert st2["requests"] == st["requests"] and st2["launches"] == st2["requests"]
    assert srv2.check_alone() == 0
    kl = _kept_results(srv2)
    # the same trace indices are kept in both modes (completion order differs)
    key = lambda t: t[0]["id"]
    assert len(kb) == len(kl)
    for (rb, cb), (rl, cl) in zip(sorted(kb, key=key), sorted(kl, key=key)):
        assert (rb["id"], rb["M"], rb["N"], rb["K"]) == (rl["id"], rl["M"], rl["N"], rl["K"])
        assert np.array_equal(cb.view(np.uint32), cl.view(np.uint32)), rb
    srv2.close()


def test_serve_tight_budget_drains(ctx):
    cfg = dict(SMALL, rate=2000.0, duration=0.05, keep=3, keep_max=20)
    # largest possible footprint: GEMV 4096 x 4096 (+x, y), GEMM 512^3 is smaller
    cfg["hbm_budget"] = 4 * (4096 * 4096 + 2 * 4096) + (1 << 16)
    _say("serve tight budget")
    srv = Server(ctx, **cfg)
    st = srv.run()
    print("[serve tight]", st, flush=True)
    assert st["peak_footprint"] <= cfg["hbm_budget"]
    assert srv.check_alone() == 0
    srv.close()


def test_serve_rejects_request_larger_than_budget(ctx):
    with pytest.raises(_capi.GatError):
        Server(ctx, **dict(SMALL, hbm_budget=1 << 20))
