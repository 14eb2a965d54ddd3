"""Mixed-precision matrix-core GEMMs on the GPU: BF16 inputs, split-BF16x3
(FP32 inputs) and FP8 e4m3 with per-128x128-block scales.

Bars (componentwise error max|C - C64| / (|alpha||op(A)||op(B)| + |beta||C0|)):
  * BF16 and FP8 kernels against the FP64 product of the operands they were
    given (bf16-rounded / dequantised values, exactly representable): <= 1e-5
    -- only FP32 accumulation separates them;
  * split-BF16x3 against the FP64 product of the original FP32 inputs: <= 1e-5;
  * the conversions themselves: BF16 rounding against torch's bfloat16 cast,
    FP8 encoding against torch's float8_e4m3fn cast of x / scale;
  * the reference's all-ones test pattern gives C == K exactly in every form.
The error of BF16 / FP8 against the original FP32 inputs is printed.
"""
import numpy as np
import pytest
import torch

from gatlas_amd import _capi
from gatlas_amd.device import Context
from oracle import oracle as ora
from tests.helpers import bf16_bits_to_f32, dequant_fp8, op, uniform_host

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def test_bf16_conversion_matches_torch(ctx):
    x = uniform_host(100003, 21, 0) * np.float32(37.5)
    dx = ctx.to_device(x)
    d16 = ctx.empty(x.size, np.uint16)
    ctx.convert_bf16(dx, d16)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(d16.download(), want)


def test_fp8_quantizer_matches_torch_cast(ctx):
    rows, cols = 300, 260
    x = uniform_host(rows * cols, 22, 0) * np.float32(3.0)
    x[5] = 0.0
    dx = ctx.to_device(x)
    q = ctx.empty(rows * cols, np.uint8)
    nb = ((rows + 127) // 128) * ((cols + 127) // 128)
    s = ctx.empty(nb, np.float32)
    ctx.quantize_fp8(dx, rows, cols, cols, False, q, s)
    qh, sh = q.download(), s.download()
    # scales: smallest power of two >= amax/448 per block
    X = x.reshape(rows, cols)
    cb = (cols + 127) // 128
    for b in range(nb):
        r0, c0 = (b // cb) * 128, (b % cb) * 128
        amax = np.abs(X[r0:r0 + 128, c0:c0 + 128]).max()
        t = np.float32(amax) / np.float32(448.0)
        want = 2.0 ** np.ceil(np.log2(np.float64(t)))
        assert sh[b] == want, (b, sh[b], want)
    srep = np.repeat(np.repeat(sh.reshape(-1, cb), 128, axis=0)[:rows], 128, axis=1)[:, :cols]
    ref = torch.from_numpy((X / srep).astype(np.float32)).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(qh.reshape(rows, cols), ref)
    # transposed read gives the transposed layout
    qt = ctx.empty(rows * cols, np.uint8)
    st = ctx.empty(nb, np.float32)
    ctx.quantize_fp8(dx, cols, rows, cols, True, qt, st)
    assert np.array_equal(qt.download().reshape(cols, rows), qh.reshape(rows, cols).T)


def test_fp8_quantizer_short_leading_dimension_rejected(ctx):
    x = ctx.to_device(uniform_host(64 * 64, 24, 0))
    q, s = ctx.empty(64 * 64, np.uint8), ctx.empty(1, np.float32)
    for trans in (0, 1):
        rc = ctx.lib.gat_quantize_fp8(ctx.h, x.ptr, 64, 64, 63, trans, q.ptr, s.ptr)
        assert rc == _capi.GAT_ERR_INVALID, (trans, rc)


@pytest.mark.parametrize("rows,cols,ld", [(131, 257, 257), (129, 200, 203), (256, 384, 388)])
def test_fp8_quantizer_ragged_and_strided(ctx, rows, cols, ld):
    """Row lengths not a multiple of 4 (element loads, single-byte stores),
    leading dimensions other than the row length, aligned or not, both read
    directions: the bytes and scales of torch's float8_e4m3fn cast of x /
    scale, and the transposed read gives the transposed layout."""
    x = uniform_host(rows * ld, 23, 0) * np.float32(5.0)
    x[7] = -0.0
    X = x.reshape(rows, ld)[:, :cols]
    dx = ctx.to_device(x)
    cb = (cols + 127) // 128
    nb = ((rows + 127) // 128) * cb
    q, s = ctx.empty(rows * cols, np.uint8), ctx.empty(nb, np.float32)
    ctx.quantize_fp8(dx, rows, cols, ld, False, q, s)
    sh = s.download()
    srep = np.repeat(np.repeat(sh.reshape(-1, cb), 128, axis=0)[:rows], 128, axis=1)[:, :cols]
    ref = torch.from_numpy((X / srep).astype(np.float32)).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(q.download().reshape(rows, cols), ref)
    # the same matrix read transposed: element (c, r) of a cols x rows operand
    rb = (rows + 127) // 128
    qt, st = ctx.empty(rows * cols, np.uint8), ctx.empty(nb, np.float32)
    ctx.quantize_fp8(dx, cols, rows, ld, True, qt, st)
    assert np.array_equal(qt.download().reshape(cols, rows), ref.T)
    assert np.array_equal(st.download().reshape(cb, rb), sh.reshape(rb, cb).T)


def _operands(M, N, K, ta, tb, seed):
    A = uniform_host(M * K, seed, 0)
    B = uniform_host(K * N, seed, 1)
    C0 = uniform_host(M * N, seed, 2)
    return A, B, C0


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(256, 256, 256), (130, 70, 200), (512, 384, 96), (1, 1, 16)])
def test_bf16x3_accuracy(ctx, ta, tb, M, N, K):
    _say("bf16x3", ta, tb, M, N, K)
    A, B, C0 = _operands(M, N, K, ta, tb, 31)
    dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.to_device(C0)
    ctx.gemm(dA, dB, dC, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, dtype=_capi.GAT_BF16X3)
    C64, MAG = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, C0=C0)
    cw, nw, _ = ora.error_metrics(dC.download(), C64, MAG)
    _say("bf16x3 err componentwise %.3e normwise %.3e" % (cw, nw))
    assert cw <= TOL, (cw, nw)


@pytest.mark.parametrize("dt", [_capi.GAT_BF16, _capi.GAT_BF16X3])
def test_bf16_every_cfg(ctx, dt):
    M, N, K = 264, 200, 144
    A, B, C0 = _operands(M, N, K, 0, 0, 32)
    for cfg in ctx.tile_cfgs(dt):
        for ta, tb in [(0, 0), (1, 1)]:
            _say("bf16 cfg", dt, cfg, ta, tb)
            dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.empty(M * N)
            if dt == _capi.GAT_BF16:
                a16, b16 = ctx.empty(A.size, np.uint16), ctx.empty(B.size, np.uint16)
                ctx.convert_bf16(dA, a16)
                ctx.convert_bf16(dB, b16)
                Ar, Br = bf16_bits_to_f32(a16.download()), bf16_bits_to_f32(b16.download())
                ctx.gemm(a16, b16, dC, M, N, K, transA=ta, transB=tb, dtype=dt, cfg=cfg)
            else:
                Ar, Br = A, B
                ctx.gemm(dA, dB, dC, M, N, K, transA=ta, transB=tb, dtype=dt, cfg=cfg)
            C64, MAG = ora.gemm64(Ar, Br, M, N, K, transA=ta, transB=tb)
            cw, _, _ = ora.error_metrics(dC.download(), C64, MAG)
            assert cw <= TOL, (dt, cfg, ta, tb, cw)


def _pad_rows(x, rows, cols, extra):
    """rows x cols row-major array -> rows x (cols + extra), zero padding."""
    out = np.zeros((rows, cols + extra), dtype=x.dtype)
    out[:, :cols] = x.reshape(rows, cols)
    return out


@pytest.mark.parametrize("K", [192, 200, 208])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_bf16_direct_lds_path(ctx, ta, tb, K):
    # M, N with whole and partial 256-wide tiles: the interior tiles of the
    # 64-deep configurations stage operands straight from HBM into LDS
    # (k-contiguous operands XOR-swizzled, row-contiguous ones read back
    # with the transposing LDS read), the edge tiles through registers;
    # K = 200 adds a short first stage of 8 k values staged through
    # registers (K = 208: a short first stage of 16).  Checked against the
    # FP64 oracle, and for K a multiple of 16 bit for bit against the
    # register path on the same data (leading dimensions one element longer,
    # not 16-byte multiples, force that path on every tile; for K = 8 mod 16
    # the two paths pad the short stage at different ends and so add the k
    # values in different groupings).
    M, N = 512, 384
    A, B, C0 = _operands(M, N, K, ta, tb, 35)
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    a16, b16 = ctx.empty(A.size, np.uint16), ctx.empty(B.size, np.uint16)
    ctx.convert_bf16(dA, a16)
    ctx.convert_bf16(dB, b16)
    ha, hb = a16.download(), b16.download()
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    pa, pb = ctx.to_device(_pad_rows(ha, ra, ca, 1)), ctx.to_device(_pad_rows(hb, rb, cb, 1))
    Ar, Br = bf16_bits_to_f32(ha), bf16_bits_to_f32(hb)
    C64, MAG = ora.gemm64(Ar, Br, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, C0=C0)
    for cfg in ctx.tile_cfgs(_capi.GAT_BF16):
        if cfg[4] != 64:
            continue
        _say("bf16 direct-LDS", cfg, ta, tb)
        d1, d2 = ctx.to_device(C0), ctx.to_device(C0)
        ctx.gemm(a16, b16, d1, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, dtype=_capi.GAT_BF16,
                 cfg=cfg)
        ctx.gemm(pa, pb, d2, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, dtype=_capi.GAT_BF16,
                 lda=ca + 1, ldb=cb + 1, cfg=cfg)
        got, ref = d1.download(), d2.download()
        cw, _, _ = ora.error_metrics(got, C64, MAG)
        assert cw <= TOL, (cfg, ta, tb, cw)
        cw2, _, _ = ora.error_metrics(ref, C64, MAG)
        assert cw2 <= TOL, (cfg, ta, tb, cw2)
        if K % 16 == 0:
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (cfg, ta, tb)


@pytest.mark.parametrize("K", [96, 112, 100])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_bf16x3_direct_lds_path(ctx, ta, tb, K):
    # Split-BF16x3 tiles of at most 2 x 2 blocks per wave stage interior
    # tiles' FP32 operands straight from HBM into LDS and split each
    # fragment as it is read; edge tiles split while staging through
    # registers.  M, N give whole and partial tiles; K = 112 adds a short
    # first stage of 16 k values staged through registers; K = 100 is not a
    # multiple of 16, so every tile takes the register path.  Checked against
    # the FP64 oracle, and bit for bit against the register path on the same
    # data (leading dimensions one element longer, not 16-byte multiples,
    # force that path on every tile).
    M, N = 512, 384
    A, B, C0 = _operands(M, N, K, ta, tb, 38)
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    pA, pB = ctx.to_device(_pad_rows(A, ra, ca, 1)), ctx.to_device(_pad_rows(B, rb, cb, 1))
    # rows padded to a 16-byte multiple: still the direct-to-LDS path, with
    # leading dimensions other than the row length
    qA, qB = ctx.to_device(_pad_rows(A, ra, ca, 4)), ctx.to_device(_pad_rows(B, rb, cb, 4))
    C64, MAG = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, C0=C0)
    for cfg in ctx.tile_cfgs(_capi.GAT_BF16X3):
        if tuple(cfg[:6]) == SPLIT_X3:
            continue
        _say("bf16x3 direct-LDS", cfg, ta, tb, K)
        d1, d2, d3 = ctx.to_device(C0), ctx.to_device(C0), ctx.to_device(C0)
        ctx.gemm(dA, dB, d1, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, dtype=_capi.GAT_BF16X3, cfg=cfg)
        ctx.gemm(pA, pB, d2, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, dtype=_capi.GAT_BF16X3,
                 lda=ca + 1, ldb=cb + 1, cfg=cfg)
        ctx.gemm(qA, qB, d3, M, N, K, transA=ta, transB=tb, alpha=1.25, beta=0.5, dtype=_capi.GAT_BF16X3,
                 lda=ca + 4, ldb=cb + 4, cfg=cfg)
        got, ref, got4 = d1.download(), d2.download(), d3.download()
        cw, _, _ = ora.error_metrics(got, C64, MAG)
        assert cw <= TOL, (cfg, ta, tb, K, cw)
        cw2, _, _ = ora.error_metrics(ref, C64, MAG)
        assert cw2 <= TOL, (cfg, ta, tb, K, cw2)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (cfg, ta, tb, K)
        assert np.array_equal(got4.view(np.uint32), ref.view(np.uint32)), (cfg, ta, tb, K, "padded rows")


SPLIT_X3 = (2, 4, 4, 2, 32, 32)  # split-plane form of split-BF16x3
REG_X3 = (2, 4, 4, 2, 16, 32)    # register-staged form, same tile


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_bf16x3_split_planes_match_register_form(ctx, ta, tb):
    """The split-plane form (operands split once into padded hi/lo planes,
    direct-to-LDS GEMM) issues the register-staged form's products in the
    same order: equal results, ragged M, N and K not a multiple of 32, in
    the general form; both within 1e-5 of the FP64 product."""
    M, N, K = 300, 520, 200
    _say("bf16x3 split planes", ta, tb, M, N, K)
    A, B, C0 = _operands(M, N, K, ta, tb, 37)
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    d1, d2 = ctx.to_device(C0), ctx.to_device(C0)
    ctx.gemm(dA, dB, d1, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, dtype=_capi.GAT_BF16X3, cfg=SPLIT_X3)
    ctx.gemm(dA, dB, d2, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, dtype=_capi.GAT_BF16X3, cfg=REG_X3)
    got, ref = d1.download(), d2.download()
    C64, MAG = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, C0=C0)
    cw, _, _ = ora.error_metrics(got, C64, MAG)
    assert cw <= TOL, cw
    assert np.array_equal(got, ref)
    # the split pass over rows padded to 16-byte multiples (vector loads) and
    # by one element (element loads) gives the same planes
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    for pad in (4, 1):
        pA, pB = ctx.to_device(_pad_rows(A, ra, ca, pad)), ctx.to_device(_pad_rows(B, rb, cb, pad))
        d3 = ctx.to_device(C0)
        ctx.gemm(pA, pB, d3, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, dtype=_capi.GAT_BF16X3,
                 lda=ca + pad, ldb=cb + pad, cfg=SPLIT_X3)
        assert np.array_equal(d3.download(), ref), pad


def test_bf16x3_split_planes_strided_batch(ctx):
    # packedCalc-style batch through the split-plane form (tuned for the
    # bucket): every problem equals its own single launch
    import ctypes
    M, N, K, P = 260, 130, 96, 3
    A = uniform_host(P * M * K, 38, 0)
    B = uniform_host(P * K * N, 38, 1)
    dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.empty(P * M * N)
    _capi.check(ctx.lib.gat_tuned_set(ctx.h, _capi.GAT_BF16X3, 0, 0, M, N, K, ctypes.byref(_capi.TileCfg(*SPLIT_X3))),
                "tuned_set")
    try:
        ctx.gemm_strided(dA, dB, dC, M, N, K, P, dtype=_capi.GAT_BF16X3)
        got = dC.download().reshape(P, M * N)
    finally:
        ctx.tuned_clear()
    for p in range(P):
        one = ctx.empty(M * N)
        ctx.gemm(ctx.to_device(A[p * M * K:(p + 1) * M * K]), ctx.to_device(B[p * K * N:(p + 1) * K * N]), one, M, N,
                 K, dtype=_capi.GAT_BF16X3, cfg=SPLIT_X3)
        assert np.array_equal(got[p], one.download()), p


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_bf16_inputs_exact_product(ctx, ta, tb):
    M, N, K = 384, 256, 512
    A, B, C0 = _operands(M, N, K, ta, tb, 33)
    dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.empty(M * N)
    a16, b16 = ctx.empty(A.size, np.uint16), ctx.empty(B.size, np.uint16)
    ctx.convert_bf16(dA, a16)
    ctx.convert_bf16(dB, b16)
    ctx.gemm(a16, b16, dC, M, N, K, transA=ta, transB=tb, dtype=_capi.GAT_BF16)
    got = dC.download()
    Ar, Br = bf16_bits_to_f32(a16.download()), bf16_bits_to_f32(b16.download())
    C64, MAG = ora.gemm64(Ar, Br, M, N, K, transA=ta, transB=tb)
    cw, nw, _ = ora.error_metrics(got, C64, MAG)
    assert cw <= TOL, cw
    C64o, MAGo = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb)
    cwo, nwo, _ = ora.error_metrics(got, C64o, MAGo)
    _say("bf16 vs original FP32 inputs: componentwise %.3e normwise %.3e" % (cwo, nwo))
    assert cwo <= 2 ** -7


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(256, 256, 256), (200, 130, 300), (128, 640, 1024)])
def test_fp8_exact_product_of_dequantised(ctx, ta, tb, M, N, K):
    _say("fp8", ta, tb, M, N, K)
    A, B, C0 = _operands(M, N, K, ta, tb, 34)
    dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.to_device(C0)
    kb = (K + 127) // 128
    A8, B8 = ctx.empty(M * K, np.uint8), ctx.empty(N * K, np.uint8)
    sA, sB = ctx.empty(((M + 127) // 128) * kb, np.float32), ctx.empty(((N + 127) // 128) * kb, np.float32)
    # op(A) as M x K and op(B)^T as N x K
    ctx.quantize_fp8(dA, M, K, M if ta else K, ta, A8, sA)
    ctx.quantize_fp8(dB, N, K, K if tb else N, not tb, B8, sB)
    ctx.gemm_fp8(A8, sA, B8, sB, dC, M, N, K, alpha=1.5, beta=-0.5)
    got = dC.download()
    Ad = dequant_fp8(A8.download(), sA.download(), M, K)          # op(A)
    Bd = dequant_fp8(B8.download(), sB.download(), N, K)          # op(B)^T
    C64 = 1.5 * Ad @ Bd.T - 0.5 * C0.reshape(M, N).astype(np.float64)
    MAG = 1.5 * np.abs(Ad) @ np.abs(Bd).T + 0.5 * np.abs(C0.reshape(M, N))
    cw, _, _ = ora.error_metrics(got, C64, MAG)
    assert cw <= TOL, cw
    # same through gat_gemm(dtype=FP8) from the FP32 operands
    dC2 = ctx.to_device(C0)
    ctx.gemm(dA, dB, dC2, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, dtype=_capi.GAT_FP8)
    assert np.array_equal(dC2.download(), got)
    C64o, MAGo = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, C0=C0)
    cwo, nwo, _ = ora.error_metrics(got, C64o, MAGo)
    _say("fp8 vs original FP32 inputs: componentwise %.3e normwise %.3e" % (cwo, nwo))
    assert cwo <= 0.25


@pytest.mark.parametrize("cfg", Context.tile_cfgs(_capi.GAT_FP8), ids=str)
@pytest.mark.parametrize("M,N,K", [(300, 520, 384), (512, 256, 200)])
def test_fp8_every_cfg(ctx, cfg, M, N, K):
    """Every FP8 tile shape (through gat_gemm_cfg from FP32 operands, ragged
    edges, K not a multiple of the 128-deep scale block) against the FP64
    product of the dequantised operands."""
    _say("fp8 cfg", cfg, M, N, K)
    A, B, C0 = _operands(M, N, K, 0, 0, 35)
    dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.to_device(C0)
    kb = (K + 127) // 128
    A8, B8 = ctx.empty(M * K, np.uint8), ctx.empty(N * K, np.uint8)
    sA, sB = ctx.empty(((M + 127) // 128) * kb, np.float32), ctx.empty(((N + 127) // 128) * kb, np.float32)
    ctx.quantize_fp8(dA, M, K, K, False, A8, sA)
    ctx.quantize_fp8(dB, N, K, N, True, B8, sB)
    ctx.gemm(dA, dB, dC, M, N, K, alpha=1.5, beta=-0.5, dtype=_capi.GAT_FP8, cfg=cfg)
    Ad = dequant_fp8(A8.download(), sA.download(), M, K)
    Bd = dequant_fp8(B8.download(), sB.download(), N, K)
    C64 = 1.5 * Ad @ Bd.T - 0.5 * C0.reshape(M, N).astype(np.float64)
    MAG = 1.5 * np.abs(Ad) @ np.abs(Bd).T + 0.5 * np.abs(C0.reshape(M, N))
    cw, _, _ = ora.error_metrics(dC.download(), C64, MAG)
    assert cw <= TOL, (cfg, cw)


class _Ptr:
    """A raw device address standing in for a DeviceArray argument."""

    def __init__(self, ptr):
        self.ptr = ptr


@pytest.mark.parametrize("cfg", Context.tile_cfgs(_capi.GAT_FP8), ids=str)
def test_fp8_direct_lds_path(ctx, cfg):
    """Resident FP8 operands, K a multiple of 128, ragged M and N: whole tiles
    run in the direct-to-LDS kernel, edge tiles in the register-staged one.
    The same bytes placed one byte past 16-byte alignment send every tile
    through the register-staged kernel; both results agree bit for bit and
    with the FP64 product of the dequantised operands."""
    import ctypes
    M, N, K = 300, 520, 384
    _say("fp8 direct-LDS", cfg, M, N, K)
    A, B, C0 = _operands(M, N, K, 0, 0, 36)
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    kb = K // 128
    A8, B8 = ctx.empty(M * K, np.uint8), ctx.empty(N * K, np.uint8)
    sA, sB = ctx.empty(((M + 127) // 128) * kb, np.float32), ctx.empty(((N + 127) // 128) * kb, np.float32)
    ctx.quantize_fp8(dA, M, K, K, False, A8, sA)
    ctx.quantize_fp8(dB, N, K, N, True, B8, sB)
    A8s, B8s = ctx.empty(M * K + 16, np.uint8), ctx.empty(N * K + 16, np.uint8)
    _capi.check(ctx.lib.gat_memcpy_d2d(ctx.h, A8s.ptr + 1, A8.ptr, M * K), "d2d")
    _capi.check(ctx.lib.gat_memcpy_d2d(ctx.h, B8s.ptr + 1, B8.ptr, N * K), "d2d")
    d1, d2 = ctx.to_device(C0), ctx.to_device(C0)
    # resident FP8 GEMMs use the tuned tile of their (FP8, NT) size bucket
    _capi.check(ctx.lib.gat_tuned_set(ctx.h, _capi.GAT_FP8, 0, 1, M, N, K, ctypes.byref(_capi.TileCfg(*cfg))),
                "tuned_set")
    try:
        ctx.gemm_fp8(A8, sA, B8, sB, d1, M, N, K, alpha=1.5, beta=-0.5)
        ctx.gemm_fp8(_Ptr(A8s.ptr + 1), sA, _Ptr(B8s.ptr + 1), sB, d2, M, N, K, alpha=1.5, beta=-0.5)
        got, ref = d1.download(), d2.download()
    finally:
        ctx.tuned_clear()
    Ad = dequant_fp8(A8.download(), sA.download(), M, K)
    Bd = dequant_fp8(B8.download(), sB.download(), N, K)
    C64 = 1.5 * Ad @ Bd.T - 0.5 * C0.reshape(M, N).astype(np.float64)
    MAG = 1.5 * np.abs(Ad) @ np.abs(Bd).T + 0.5 * np.abs(C0.reshape(M, N))
    cw, _, _ = ora.error_metrics(got, C64, MAG)
    assert cw <= TOL, (cfg, cw)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), cfg


@pytest.mark.parametrize("dt", [_capi.GAT_BF16X3, _capi.GAT_FP8, _capi.GAT_BF16])
def test_mixed_test_pattern(ctx, dt):
    M = N = K = 1024
    ones = ctx.to_device(np.ones(M * K, np.float32))
    dC = ctx.empty(M * N)
    if dt == _capi.GAT_BF16:
        o16 = ctx.empty(M * K, np.uint16)
        ctx.convert_bf16(ones, o16)
        ctx.gemm(o16, o16, dC, M, N, K, dtype=dt)
    else:
        ctx.gemm(ones, ones, dC, M, N, K, dtype=dt)
    assert ora.check_array(dC.download(), float(K))


@pytest.mark.parametrize("dist", ["normal", "illscaled"])
def test_bf16x3_8192_sampled(ctx, dist):
    # Mixed-precision config: M=N=K=8192, A, B ~ N(0,1); "illscaled"
    # multiplies column k of A and column j of B by 10^u, u ~ U(-3, 3).
    n = 8192
    _say("bf16x3 8192", dist)
    dA, dB, dC = ctx.empty(n * n), ctx.empty(n * n), ctx.empty(n * n)
    ctx.fill_normal(dA, 0x4D495845, 0)
    ctx.fill_normal(dB, 0x4D495845, 1)
    if dist == "illscaled":
        ctx.scale_columns(dA, n, n, 0x4D495845, 10)
        ctx.scale_columns(dB, n, n, 0x4D495845, 11)
    ctx.gemm(dA, dB, dC, n, n, n, dtype=_capi.GAT_BF16X3)
    rows = np.arange(0, n, 409)
    got = dC.download().reshape(n, n)[rows]
    A, B = dA.download(), dB.download()
    C64, MAG = ora.gemm64(A, B, n, n, n, rows=rows)
    cw, nw, _ = ora.error_metrics(got.ravel(), C64, MAG)
    _say("bf16x3 8192 %s: componentwise %.3e normwise %.3e" % (dist, cw, nw))
    assert cw <= TOL
    for a in (dA, dB, dC):
        a.free()


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("dist", ["normal", "illscaled"])
def test_bf16x3_2048_direct_lds_sampled(ctx, dist, ta, tb):
    # 2048^3 takes the mid-size default: 128 x 128 tiles, every tile interior,
    # on the direct-to-LDS x3 routine; general form, row sample against FP64.
    n = 2048
    _say("bf16x3 2048", dist, ta, tb)
    dA, dB = ctx.empty(n * n), ctx.empty(n * n)
    ctx.fill_normal(dA, 0x4D495846, 0)
    ctx.fill_normal(dB, 0x4D495846, 1)
    if dist == "illscaled":
        ctx.scale_columns(dA, n, n, 0x4D495846, 10)
        ctx.scale_columns(dB, n, n, 0x4D495846, 11)
    C0 = uniform_host(n * n, 39, 2)
    dC = ctx.to_device(C0)
    ctx.gemm(dA, dB, dC, n, n, n, bool(ta), bool(tb), alpha=1.5, beta=-0.5, dtype=_capi.GAT_BF16X3)
    rows = np.arange(0, n, 97)
    got = dC.download().reshape(n, n)[rows]
    A, B = dA.download(), dB.download()
    C64, MAG = ora.gemm64(A, B, n, n, n, transA=ta, transB=tb, alpha=1.5, beta=-0.5, C0=C0, rows=rows)
    cw, nw, _ = ora.error_metrics(got.ravel(), C64, MAG)
    _say("bf16x3 2048 %s %d%d: componentwise %.3e normwise %.3e" % (dist, ta, tb, cw, nw))
    assert cw <= TOL


def test_fp8_resident_without_scales_rejected(ctx):
    # resident e4m3 operands without block scales: rejected on the host before
    # any launch, C untouched
    import ctypes
    M = N = K = 128
    A8, B8 = ctx.empty(M * K, np.uint8), ctx.empty(N * K, np.uint8)
    s = ctx.empty(1, np.float32)
    C0 = uniform_host(M * N, 38, 0)
    dC = ctx.to_device(C0)
    for sa, sb in [(None, s.ptr), (s.ptr, None), (None, None)]:
        rc = ctx.lib.gat_gemm_fp8(ctx.h, M, N, K, 1.0, A8.ptr, ctypes.c_void_p(sa), B8.ptr, ctypes.c_void_p(sb),
                                  0.0, dC.ptr, N)
        assert rc == _capi.GAT_ERR_INVALID, (sa, sb, rc)
    ctx.sync()
    assert np.array_equal(dC.download().view(np.uint32), C0.view(np.uint32))


def test_gemm_batch_above_int32_rejected(ctx):
    # the batch count is checked before it is narrowed: 2^32 + 1 problems is
    # an error, not one problem
    M = N = K = 16
    dA, dB, dC = ctx.empty(M * K), ctx.empty(K * N), ctx.empty(M * N)
    rc = ctx.lib.gat_gemm_strided(ctx.h, _capi.GAT_F32, 0, 0, M, N, K, 1.0, dA.ptr, K, 0, dB.ptr, N, 0, 0.0,
                                  dC.ptr, N, 0, (1 << 32) + 1)
    assert rc == _capi.GAT_ERR_INVALID
    assert "batch" in _capi.last_error()


def test_fp8_instruction_truncation_through_gemm(ctx):
    """Crafted resident e4m3 operands (scales 1.0) through gat_gemm_fp8: the
    sums match the model of the matrix instruction (groups of 8 k summed in
    fixed point at 2^(E-13)), including the products it drops, which the
    exact product keeps."""
    from tests.helpers import fp8_mfma_model
    from tests.test_formats import FP8_MEASURED_VECTORS
    M = N = K = 128
    pats = [(np.array(a, np.uint8), np.array(b, np.uint8)) for a, b, _ in FP8_MEASURED_VECTORS]
    A8 = np.zeros((M, K), np.uint8)
    B8 = np.zeros((N, K), np.uint8)
    for i in range(M):
        a, _ = pats[i % len(pats)]
        A8[i, (i % 2) * 64:(i % 2) * 64 + 64] = a
    for j in range(N):
        _, b = pats[(j // 2) % len(pats)]
        B8[j, (j % 2) * 64:(j % 2) * 64 + 64] = b
    dA8, dB8 = ctx.to_device(A8.ravel()), ctx.to_device(B8.ravel())
    one = ctx.to_device(np.ones(1, np.float32))
    dC = ctx.empty(M * N)
    ctx.gemm_fp8(dA8, one, dB8, one, dC, M, N, K)
    got = dC.download().reshape(M, N).astype(np.float64)
    model = fp8_mfma_model(A8, np.ones((M, 1)), B8, np.ones((N, 1)), K)
    assert np.allclose(got, model, rtol=2.0 ** -22, atol=0), np.max(np.abs(got - model))
    # the diagonal-pattern entries reproduce the measured single-instruction sums
    for i in range(len(pats)):
        want = FP8_MEASURED_VECTORS[i][2]
        j = 2 * i + (i % 2)      # row i's pattern meets column j's pattern at the same k half
        assert got[i, j] == want, (i, got[i, j], want)
