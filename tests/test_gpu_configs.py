"""Parity at the benchmark workloads' own sizes (called through the C ABI).

Workloads and bars:
  * plain sweep, general form C = 1.5*op(A)*op(B) - 0.5*C, entries U[-1,1]:
      - FP64 at 1024^3 and 2048^3, every transpose pair, every output checked
        against the FP64 oracle: componentwise error <= 1e-13;
      - FP64 and FP32 at 4096^3 and 8192^3 (FP32 8192^3 in the three
        transposed pairs; NN is in test_gpu_kernels), sampled rows against
        the oracle: <= 1e-13 (FP64), <= 1e-5 (FP32);
  * SGEMV y = A*x + y at M = N = 16384, both storage orders, every output
    against the FP64 oracle: <= 1e-5; the reference-geometry GEMV at the same
    size bit-exact against the reference's own host check
    (KernelMatvecBuffer.hpp:118-133) on its paranoid drand48 inputs;
  * mixed precision at 8192^3, A, B ~ N(0,1) and the ill-scaled variant
    (column k of A and column j of B times 10^u, u ~ U(-3,3)):
      - BF16 inputs: sampled rows against the FP64 product of the
        bf16-rounded operands: <= 1e-5 (only FP32 accumulation separates
        them);
      - FP8 e4m3 with 128x128 block scales: sampled rows against a model of
        the matrix instruction's own sums (groups of 8 k summed in fixed
        point, tests.helpers.fp8_mfma_model): <= 1e-5; against the exact
        FP64 product of the dequantised operands: <= 1e-5 for N(0,1) data
        and <= 7*2^-13 + 1e-5 (the instruction's truncation bound) for the
        ill-scaled data;
    the error against the original FP32 inputs is printed for both;
  * served requests at the trace's own shapes (GEMM M, N, K in [64, 4096],
    GEMV M, N in [1024, 32768], FP32, 2k req/s, 64 GB budget): every kept
    result bit-identical to the same request run alone, and within 1e-5 of
    the FP64 oracle on sampled rows.
The componentwise error is max|C - C64| / (|alpha||op(A)||op(B)| + |beta||C0|)
(with the reference's paranoid rule OCLAppUtil.hpp:212-237
as the bit-exact bar of the reference-geometry kernels).
"""
import numpy as np
import pytest

from gatlas_amd import _capi
from gatlas_amd.device import Server
from oracle import oracle as ora
from tests.helpers import bf16_bits_to_f32, dequant_fp8, e4m3_table, fp8_mfma_model, op

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5
F64_TOL = 1e-13
SEED = 0x6A7A5EED
PAIRS = [(0, 0), (1, 0), (0, 1), (1, 1)]


def _say(*a):
    print("[case]", *a, flush=True)


def _general_form(ctx, n, ta, tb, dt):
    """Device-generated A, B, C0 ~ U[-1,1] (streams 0, 1, 2), C = 1.5 op(A)op(B) - 0.5 C0.
    Returns the host copies of A, B, C0 and the result."""
    dA, dB, dC = ctx.empty(n * n, dt), ctx.empty(n * n, dt), ctx.empty(n * n, dt)
    ctx.fill_uniform(dA, SEED, 0)
    ctx.fill_uniform(dB, SEED, 1)
    ctx.fill_uniform(dC, SEED, 2)
    A, B, C0 = dA.download(), dB.download(), dC.download()
    ctx.gemm(dA, dB, dC, n, n, n, transA=ta, transB=tb, alpha=1.5, beta=-0.5)
    got = dC.download()
    for a in (dA, dB, dC):
        a.free()
    return A, B, C0, got


def _oracle_rows(A, B, C0, n, ta, tb, rows, alpha=1.5, beta=-0.5):
    """FP64 oracle on the selected rows of op(A); op(B) made contiguous first so
    the oracle's inner loop runs along memory (same sums, same order)."""
    opA = np.ascontiguousarray(op(A, n, n, ta)[rows])
    opB = np.ascontiguousarray(op(B, n, n, tb))
    C0r = None if C0 is None else np.ascontiguousarray(C0.reshape(n, n)[rows])
    return ora.gemm64(opA, opB, len(rows), n, n, alpha=alpha, beta=beta, C0=C0r)


# ------------------------------------------------------- plain sweep, FP64
@pytest.mark.parametrize("ta,tb", PAIRS)
@pytest.mark.parametrize("n", [1024, 2048])
def test_dgemm_general_form_full(ctx, n, ta, tb):
    _say("dgemm general form full", n, ta, tb)
    A, B, C0, got = _general_form(ctx, n, ta, tb, np.float64)
    C64, MAG = _oracle_rows(A, B, C0, n, ta, tb, np.arange(n))
    cw, nw, _ = ora.error_metrics(got, C64, MAG)
    _say("dgemm %d %d%d componentwise %.3e normwise %.3e" % (n, ta, tb, cw, nw))
    assert cw <= F64_TOL, (n, ta, tb, cw)


def _sample_rows(n):
    rows = [0, 1, 63, 64, 127, 128, n // 2 - 1, n // 2, n - 2, n - 1]
    rows += list(range(1000 % n, n, max(1, n // 11)))
    return np.unique(np.array(rows, dtype=np.int64))


@pytest.mark.parametrize("ta,tb", PAIRS)
@pytest.mark.parametrize("n", [4096, 8192])
def test_dgemm_general_form_sampled(ctx, n, ta, tb):
    _say("dgemm general form sampled", n, ta, tb)
    A, B, C0, got = _general_form(ctx, n, ta, tb, np.float64)
    rows = _sample_rows(n)
    C64, MAG = _oracle_rows(A, B, C0, n, ta, tb, rows)
    cw, nw, _ = ora.error_metrics(np.ascontiguousarray(got.reshape(n, n)[rows]), C64, MAG)
    _say("dgemm %d %d%d componentwise %.3e normwise %.3e" % (n, ta, tb, cw, nw))
    assert cw <= F64_TOL, (n, ta, tb, cw)


@pytest.mark.parametrize("n,ta,tb", [(4096, 0, 0), (4096, 1, 0), (4096, 0, 1), (4096, 1, 1),
                                     (8192, 1, 0), (8192, 0, 1), (8192, 1, 1)])
def test_sgemm_general_form_sampled(ctx, n, ta, tb):
    _say("sgemm general form sampled", n, ta, tb)
    A, B, C0, got = _general_form(ctx, n, ta, tb, np.float32)
    rows = _sample_rows(n)
    C64, MAG = _oracle_rows(A, B, C0, n, ta, tb, rows)
    cw, nw, _ = ora.error_metrics(np.ascontiguousarray(got.reshape(n, n)[rows]), C64, MAG)
    _say("sgemm %d %d%d componentwise %.3e normwise %.3e" % (n, ta, tb, cw, nw))
    assert cw <= F32_TOL and nw <= F32_TOL, (n, ta, tb, cw, nw)


# ------------------------------------------------------------ SGEMV 16384^2
@pytest.mark.parametrize("ta", [0, 1])
def test_sgemv_16384_full(ctx, ta):
    M = N = 16384
    _say("sgemv 16384^2", ta)
    dA, dx, dy = ctx.empty(M * N), ctx.empty(N), ctx.empty(M)
    ctx.fill_uniform(dA, SEED, 3)
    ctx.fill_uniform(dx, SEED, 4)
    ctx.fill_uniform(dy, SEED, 5)
    A, x, y0 = dA.download(), dx.download(), dy.download()
    ctx.gemv(dA, dx, dy, M, N, transA=ta, alpha=1.0, beta=1.0)
    got = dy.download()
    y64, mag = ora.gemv64(A, x, M, N, transA=ta, alpha=1.0, beta=1.0, y0=y0)
    cw, nw, _ = ora.error_metrics(got, y64, mag)
    _say("sgemv 16384^2 ta=%d componentwise %.3e normwise %.3e" % (ta, cw, nw))
    assert cw <= F32_TOL and nw <= F32_TOL, (ta, cw, nw)
    for a in (dA, dx, dy):
        a.free()


@pytest.mark.parametrize("ta", [0, 1])
def test_sgemv_compat_16384_paranoid(ctx, ta):
    # The reference's GEMV at its benchmark size, paranoid mode: alpha, beta,
    # A, x, y from the unseeded drand48 stream, work-group 64, 8 rows per
    # work-item, float4 (KernelMatvecBuffer.hpp:81-177); bit-exact against
    # its host check, which the paranoid rule then passes.
    M = N = 16384
    _say("sgemv compat 16384^2 paranoid", ta)
    r = ora.Drand48(0)
    alpha, beta = r.posrand(np.float32), r.posrand(np.float32)
    A = r.fillrand(M * N, np.float32)
    x = r.fillrand(N, np.float32)
    y0 = r.fillrand(M, np.float32)
    want = ora.ref_matvec(A, x, M, N, transA=ta, gemv=True, alpha=alpha, beta=beta, yold=y0)
    dA, dx, dy = ctx.to_device(A), ctx.to_device(x), ctx.to_device(y0)
    ctx.gemv_compat(dA, dx, dy, M, N, 64, 8, 4, transA=ta, gemv=True, alpha=float(alpha), beta=float(beta))
    got = dy.download()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert ora.paranoid_pass(got, want)
    for a in (dA, dx, dy):
        a.free()


# ------------------------------------------------- mixed precision 8192^3
MIX_SEED = 0x4D495845


def _mixed_operands(ctx, n, dist):
    dA, dB = ctx.empty(n * n), ctx.empty(n * n)
    ctx.fill_normal(dA, MIX_SEED, 0)
    ctx.fill_normal(dB, MIX_SEED, 1)
    if dist == "illscaled":
        ctx.scale_columns(dA, n, n, MIX_SEED, 10)
        ctx.scale_columns(dB, n, n, MIX_SEED, 11)
    return dA, dB


@pytest.mark.parametrize("dist", ["normal", "illscaled"])
def test_bf16_8192_sampled(ctx, dist):
    n = 8192
    _say("bf16 8192", dist)
    dA, dB = _mixed_operands(ctx, n, dist)
    a16, b16, dC = ctx.empty(n * n, np.uint16), ctx.empty(n * n, np.uint16), ctx.empty(n * n)
    ctx.convert_bf16(dA, a16)
    ctx.convert_bf16(dB, b16)
    ctx.gemm(a16, b16, dC, n, n, n, dtype=_capi.GAT_BF16)
    rows = _sample_rows(n)
    got = np.ascontiguousarray(dC.download().reshape(n, n)[rows])
    Ar, Br = bf16_bits_to_f32(a16.download()), bf16_bits_to_f32(b16.download())
    C64, MAG = _oracle_rows(Ar, Br, None, n, 0, 0, rows, alpha=1.0, beta=0.0)
    cw, nw, _ = ora.error_metrics(got, C64, MAG)
    A, B = dA.download(), dB.download()
    C64o, MAGo = _oracle_rows(A, B, None, n, 0, 0, rows, alpha=1.0, beta=0.0)
    cwo, nwo, _ = ora.error_metrics(got, C64o, MAGo)
    _say("bf16 8192 %s: vs bf16 operands %.3e; vs FP32 inputs componentwise %.3e normwise %.3e"
         % (dist, cw, cwo, nwo))
    assert cw <= F32_TOL, (dist, cw)
    for a in (dA, dB, a16, b16, dC):
        a.free()


# One scaled FP8 matrix instruction sums each group of 8 consecutive k in
# fixed point: a product more than 13 binades below its group's largest
# exponent sum is dropped, and others are truncated toward zero to that step
# (measured; modelled by tests.helpers.fp8_mfma_model).  Against the exact
# product of the dequantised operands this adds at most 7 * 2^-13 of
# sum|a||b| per output; inputs whose products stay within 2^-13 of each other
# inside a group of 8 (normal data) see only the FP32 accumulation.
FP8_GROUP_BOUND = 7 * 2.0 ** -13


@pytest.mark.parametrize("dist", ["normal", "illscaled"])
def test_fp8_8192_sampled(ctx, dist):
    n = 8192
    _say("fp8 8192", dist)
    dA, dB = _mixed_operands(ctx, n, dist)
    kb = n // 128
    A8, B8 = ctx.empty(n * n, np.uint8), ctx.empty(n * n, np.uint8)
    sA, sB = ctx.empty(kb * kb, np.float32), ctx.empty(kb * kb, np.float32)
    # op(A) as M x K, op(B)^T as N x K (B stored K x N row-major, read transposed)
    ctx.quantize_fp8(dA, n, n, n, False, A8, sA)
    ctx.quantize_fp8(dB, n, n, n, True, B8, sB)
    dC = ctx.empty(n * n)
    ctx.gemm_fp8(A8, sA, B8, sB, dC, n, n, n)
    rows = _sample_rows(n)
    Ch = dC.download().reshape(n, n)
    got = np.ascontiguousarray(Ch[rows])
    t = e4m3_table()
    A8h, B8h = A8.download().reshape(n, n), B8.download().reshape(n, n)
    sAh, sBh = sA.download().reshape(kb, kb).astype(np.float64), sB.download().reshape(kb, kb).astype(np.float64)
    Ad = t[A8h[rows]] * np.repeat(sAh[rows // 128], 128, axis=1)
    BdT = np.ascontiguousarray(dequant_fp8(B8h.ravel(), sBh.ravel(), n, n).T)   # K x N
    C64, MAG = ora.gemm64(np.ascontiguousarray(Ad), BdT, len(rows), n, n)
    del BdT
    cw, nw, _ = ora.error_metrics(got, C64, MAG)
    # the instruction's own sums on a column sample of the sampled rows
    cols = np.unique(np.concatenate([np.arange(0, n, 16), [1, 127, 128, n - 1]]))
    model = fp8_mfma_model(A8h[rows], sAh[rows // 128], B8h[cols], sBh[cols // 128], n)
    Cs = np.ascontiguousarray(got[:, cols])
    cwm, _, _ = ora.error_metrics(Cs, model, np.ascontiguousarray(MAG[:, cols]))
    A, B = dA.download(), dB.download()
    C64o, MAGo = _oracle_rows(A, B, None, n, 0, 0, rows, alpha=1.0, beta=0.0)
    cwo, nwo, _ = ora.error_metrics(got, C64o, MAGo)
    _say("fp8 8192 %s: vs instruction model %.3e; vs dequantised operands %.3e; vs FP32 inputs componentwise "
         "%.3e normwise %.3e" % (dist, cwm, cw, cwo, nwo))
    assert cwm <= F32_TOL, (dist, cwm)
    assert cw <= (F32_TOL if dist == "normal" else FP8_GROUP_BOUND + F32_TOL), (dist, cw)
    for a in (dA, dB, A8, B8, sA, sB, dC):
        a.free()


# ---------------------------------------------- served requests, own shapes
SERVED = dict(rate=2000.0, duration=0.4, gemm_min=64, gemm_max=4096, gemv_min=1024, gemv_max=32768,
              hbm_budget=64 * 10 ** 9, keep=9, keep_max=40, seed=23, gemm_dtype=_capi.GAT_F32)


def test_served_trace_config_shapes(ctx):
    _say("served trace at config shapes, 2k req/s")
    srv = Server(ctx, **SERVED)
    try:
        st = srv.run()
        print("[served]", {k: round(v, 4) if isinstance(v, float) else v for k, v in st.items()}, flush=True)
        assert st["requests"] == st["gemm_requests"] + st["gemv_requests"] > 500
        assert st["peak_footprint"] <= SERVED["hbm_budget"]
        assert st["kept"] >= 30
        assert srv.check_alone() == 0
        nk = {0: 0, 1: 0}
        for r in srv.kept():
            M, N, K = r["M"], r["N"], r["K"]
            assert min(M, N) >= (64 if r["kind"] == 0 else 1024)
            rows = np.unique(np.linspace(0, M - 1, min(M, 6)).astype(np.int64))
            if r["kind"] == 0:
                ta, tb = bool(r["transA"]), bool(r["transB"])
                A = srv.read(r["A"], M * K)          # lda == M (ta) or K: dense
                B = srv.read(r["B"], K * N)
                opA = np.ascontiguousarray(op(A, M, K, ta)[rows])
                opB = np.ascontiguousarray(op(B, K, N, tb))
                ref, mag = ora.gemm64(opA, opB, len(rows), N, K)
                got = srv.read(r["C"], M * N).reshape(M, N)[rows]
            else:
                lda = r["lda"]
                A = np.concatenate([srv.read(r["A"] + 4 * int(i) * lda, N) for i in rows])
                x = srv.read(r["B"], N)
                ref, mag = ora.gemv64(A, x, len(rows), N)
                got = srv.read(r["C"], M)[rows]
            cw, _, _ = ora.error_metrics(np.ascontiguousarray(got), ref, mag)
            assert cw <= F32_TOL, (r, cw)
            nk[r["kind"]] += 1
        _say("served kept checked: %d GEMM, %d GEMV" % (nk[0], nk[1]))
        assert nk[0] > 0 and nk[1] > 0
    finally:
        srv.close()
