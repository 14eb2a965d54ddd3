"""Host-only checks of the grouped GEMM's schedule (gat_grouped_plan, the
tile list gat_grouped_create launches), no device needed:

  * every element of every problem's output is covered by exactly one tile,
    and no tile starts outside its problem;
  * tile shapes are the form's: 32-wide strips (shapes 4, 5) only in the FP32
    and BF16 forms and only beside a 128-wide side, merged 256-wide tiles
    (6, 7) only in the BF16 form, only for interior shape 128x128, K a
    multiple of 16 and 16-byte aligned rows, and only wholly inside the
    problem;
  * the given interior shapes are used, and the count matches with a short
    output buffer.
"""
import ctypes

import numpy as np
import pytest

from gatlas_amd import _capi

@pytest.fixture(scope="module")
def lib():
    return _capi.load()


BM = [128, 64, 128, 64, 128, 32, 256, 128]
BN = [128, 64, 64, 128, 32, 128, 128, 256]
FORMS = [(_capi.GAT_F32, 4), (_capi.GAT_BF16, 2), (_capi.GAT_BF16X3, 4)]


def _descs(dims, esize, aligned=True):
    arr = (_capi.GemmDesc * len(dims))()
    for i, (M, N, K, ta, tb) in enumerate(dims):
        # fake, never dereferenced device addresses: 16-byte aligned or not
        base = 0x100000 * (i + 1) + (0 if aligned else esize)
        lda = (M if ta else K)
        ldb = (K if tb else N)
        arr[i] = _capi.GemmDesc(ta, tb, M, N, K, 1.0, base, lda, base + 0x80000, ldb, 0.0, base + 0x40000, N)
    return arr


def _plan(lib, dtype, dims, shapes=None, aligned=True, esize=4):
    descs = _descs(dims, esize, aligned)
    sh = None if shapes is None else (ctypes.c_int * len(dims))(*shapes)
    n = lib.gat_grouped_plan(dtype, descs, len(dims), sh, None, 0)
    assert n >= 0, _capi.last_error()
    out = (ctypes.c_int * (4 * max(1, n)))()
    assert lib.gat_grouped_plan(dtype, descs, len(dims), sh, out, n) == n
    return np.frombuffer(out, dtype=np.int32)[:4 * n].reshape(n, 4)


def _check(dtype, dims, tiles, aligned):
    cover = [np.zeros((M, N), np.int32) for M, N, _, _, _ in dims]
    for p, m0, n0, s in tiles:
        M, N, K, ta, tb = dims[p]
        assert 0 <= s < 8 and 0 <= m0 < M and 0 <= n0 < N, (p, m0, n0, s)
        if s in (4, 5):
            assert dtype != _capi.GAT_BF16X3
        if s in (6, 7):
            e = 8 if dtype == _capi.GAT_BF16 else 4
            assert dtype == _capi.GAT_BF16 and K >= 32 and K % 16 == 0 and aligned, (dims[p], s)
            assert m0 + BM[s] <= M and n0 + BN[s] <= N, (dims[p], m0, n0, s)
        cover[p][m0:m0 + BM[s], n0:n0 + BN[s]] += 1
    for (M, N, _, _, _), c in zip(dims, cover):
        assert c.min() == 1 and c.max() == 1, (M, N, c.min(), c.max())


@pytest.mark.parametrize("dtype,esize", FORMS, ids=["f32", "bf16", "bf16x3"])
@pytest.mark.parametrize("aligned", [True, False], ids=["aligned", "unaligned"])
def test_plan_covers_every_output_once(lib, dtype, esize, aligned):
    rng = np.random.default_rng(41)
    v = np.clip(np.round(np.exp(rng.uniform(np.log(16), np.log(1500), (60, 3))) / 16) * 16, 16, 1500).astype(int)
    t = rng.integers(0, 2, (60, 2))
    dims = [(int(a), int(b), int(c), int(x), int(y)) for (a, b, c), (x, y) in zip(v, t)]
    dims += [(1, 1, 1, 0, 0), (257, 33, 40, 1, 0), (640, 640, 48, 0, 1), (384, 1024, 24, 1, 1), (96, 160, 16, 0, 0)]
    tiles = _plan(lib, dtype, dims, aligned=aligned, esize=esize)
    _check(dtype, dims, tiles, aligned)
    merged = int(np.isin(tiles[:, 3], (6, 7)).sum())
    if dtype == _capi.GAT_BF16 and aligned:
        assert merged > 0
    else:
        assert merged == 0


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_plan_uses_given_shapes(lib, shape):
    dims = [(512, 384, 64, 0, 0), (200, 96, 32, 1, 1)]
    tiles = _plan(lib, _capi.GAT_F32, dims, shapes=[shape, shape])
    _check(_capi.GAT_F32, dims, tiles, True)
    # every whole tile of the first problem has the given interior shape
    first = tiles[tiles[:, 0] == 0]
    whole = [s for _, m0, n0, s in first if m0 + BM[shape] <= 512 and n0 + BN[shape] <= 384]
    assert whole and all(s == shape for s in whole)


def test_plan_short_buffer_and_errors(lib):
    dims = [(300, 300, 64, 0, 0)]
    descs = _descs(dims, 4)
    n = lib.gat_grouped_plan(_capi.GAT_F32, descs, 1, None, None, 0)
    out = (ctypes.c_int * 8)()
    assert lib.gat_grouped_plan(_capi.GAT_F32, descs, 1, None, out, 2) == n > 2
    bad = (ctypes.c_int * 1)(9)
    assert lib.gat_grouped_plan(_capi.GAT_F32, descs, 1, bad, None, 0) == _capi.GAT_ERR_INVALID
    assert lib.gat_grouped_plan(_capi.GAT_FP8, descs, 1, None, None, 0) == _capi.GAT_ERR_INVALID
