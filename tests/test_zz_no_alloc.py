"""The allocation-free GEMV entry point (gat_gemv_noalloc) through the C ABI.

  * its result is bit-identical to gat_gemv's on the same inputs, in both
    storage orders and both precisions, on two calls in a row (the second
    smaller, so the static scratch holds the first call's partials when the
    second starts), and within 1e-5 of the FP64 oracle on the smaller one;
  * a transposed shape whose slice partials exceed the static scratch is
    rejected on the host with GAT_ERR_INVALID before any launch, and y is
    left untouched.
"""
import numpy as np
import pytest

from gatlas_amd import _capi
from oracle import oracle as ora

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _gemv(ctx, fn, dt, ta, M, N, A, x, y, lda):
    code = _capi.GAT_F32 if dt == np.float32 else _capi.GAT_F64
    return getattr(ctx.lib, fn)(ctx.h, code, ta, M, N, 1.5, A.ptr, lda, x.ptr, -0.5, y.ptr)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("ta", [1, 0])
def test_gemv_noalloc_matches_gemv(ctx, ta, dt):
    uint = np.uint32 if dt == np.float32 else np.uint64
    for M, N in [(16384, 16384), (3000, 5001)]:
        _say("gemv_noalloc", dt.__name__, ta, M, N)
        lda = M if ta else N
        dA, dx = ctx.empty(M * N, dt), ctx.empty(N, dt)
        y_ref, y_na = ctx.empty(M, dt), ctx.empty(M, dt)
        ctx.fill_uniform(dA, 71, 0)
        ctx.fill_uniform(dx, 71, 1)
        ctx.fill_uniform(y_ref, 71, 2)
        ctx.fill_uniform(y_na, 71, 2)
        y0 = y_ref.download()
        assert _gemv(ctx, "gat_gemv_noalloc", dt, ta, M, N, dA, dx, y_na, lda) == 0, _capi.last_error()
        assert _gemv(ctx, "gat_gemv", dt, ta, M, N, dA, dx, y_ref, lda) == 0, _capi.last_error()
        ctx.sync()
        got, ref = y_na.download(), y_ref.download()
        assert np.array_equal(got.view(uint), ref.view(uint)), (dt, ta, M, N)
        if M * N < 10 ** 8:
            A = dA.download().astype(np.float32) if dt == np.float32 else dA.download()
            if dt == np.float32:
                y64, mag = ora.gemv64(A, dx.download(), M, N, transA=ta, alpha=1.5, beta=-0.5, y0=y0)
                cw, _, _ = ora.error_metrics(got, y64, mag)
                assert cw <= 1e-5, cw
        for a in (dA, dx, y_ref, y_na):
            a.free()


def test_gemv_noalloc_rejects_partials_above_scratch(ctx):
    # transposed, 2^23 outputs: one slice of 2^23 FP32 partials = 32 MiB,
    # twice the static scratch; rejected before any launch
    M, N = 1 << 23, 4
    assert 4 * M > _capi.GAT_GEMV_NOALLOC_SCRATCH
    dA, dx, dy = ctx.empty(M * N), ctx.empty(N), ctx.empty(M)
    ctx.fill_uniform(dy, 72, 0)
    before = dy.download()
    rc = _gemv(ctx, "gat_gemv_noalloc", np.float32, 1, M, N, dA, dx, dy, M)
    assert rc == _capi.GAT_ERR_INVALID
    assert "scratch" in _capi.last_error()
    ctx.sync()
    assert np.array_equal(dy.download().view(np.uint32), before.view(np.uint32))
    for a in (dA, dx, dy):
        a.free()
