"""Persistent grouped GEMM on the GPU.

  * every problem of a mixed group (ragged sizes, all transpose pairs, general
    alpha/beta, empty problems) against the FP64 oracle: componentwise
    error <= 1e-5 (FP32 and split-BF16x3, against the FP32 operands) and
    <= 1e-5 against the bf16-rounded operands (BF16);
  * each problem's result is bit-identical to a single launch of the same
    tile configuration (a problem computes the same alone as in the group),
    also on the BF16 form's merged 256 x 128 and 128 x 256 tiles;
  * the host-only dry run (gat_grouped_plan) given a launch's interior
    shapes has as many tiles as that launch;
  * a captured graph replayed several times, with eager work between the
    replays, gives the eager result every time;
  * 512 problems of the grouped workload (M, N, K log-uniform in [64, 4096],
    multiples of 16, random transposes): bit-identical to the per-problem
    launch loop, a row sample of every problem against the FP64 oracle.
"""
import ctypes

import numpy as np
import pytest

from gatlas_amd import _capi
from oracle import oracle as ora
from tests.helpers import bf16_bits_to_f32, uniform_host

pytestmark = pytest.mark.gpu
TOL = 1e-5
# single-problem tile of each grouped shape's numerics (32x32 blocks, no
# k-split: every shape sums an element's products in the same k order, so the
# 128 x 64 shapes compare against the 128 x 128 tile)
_BIG = {_capi.GAT_F32: (2, 2, 2, 2, 32, 32), _capi.GAT_BF16: (2, 2, 2, 2, 32, 32),
        _capi.GAT_BF16X3: (2, 2, 2, 2, 16, 32)}
_SMALL = {_capi.GAT_F32: (2, 2, 1, 1, 32, 32), _capi.GAT_BF16: (2, 2, 1, 1, 64, 32),
          _capi.GAT_BF16X3: (2, 2, 1, 1, 32, 32)}
SHAPE_CFG = {dt: [_BIG[dt], _SMALL[dt], _BIG[dt], _BIG[dt]] for dt in _BIG}
FORMS = [_capi.GAT_F32, _capi.GAT_BF16, _capi.GAT_BF16X3]
FORM_IDS = ["f32", "bf16", "bf16x3"]


def _say(*a):
    print("[case]", *a, flush=True)


def _bf16_round(x):
    import torch
    return torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)


def _make_group(ctx, dims, seed, dtype=_capi.GAT_F32, general=True):
    """Host operands + device problems for a list of (M, N, K, tA, tB)."""
    host, probs = [], []
    for i, (M, N, K, ta, tb) in enumerate(dims):
        A = uniform_host(M * K, seed, 3 * i)
        B = uniform_host(K * N, seed, 3 * i + 1)
        C0 = uniform_host(M * N, seed, 3 * i + 2)
        alpha, beta = (1.5, -0.5) if general else (1.0, 0.0)
        if dtype == _capi.GAT_BF16:
            a16, b16 = _bf16_round(A), _bf16_round(B)
            dA, dB = ctx.to_device(a16), ctx.to_device(b16)
            A, B = bf16_bits_to_f32(a16), bf16_bits_to_f32(b16)
        else:
            dA, dB = ctx.to_device(A), ctx.to_device(B)
        dC = ctx.to_device(C0)
        host.append((A, B, C0, alpha, beta))
        probs.append(dict(transA=ta, transB=tb, M=M, N=N, K=K, alpha=alpha, beta=beta, A=dA, B=dB, C=dC))
    return host, probs


def _check_oracle(dims, host, probs, rows_per=None):
    worst = 0.0
    for (M, N, K, ta, tb), (A, B, C0, alpha, beta), p in zip(dims, host, probs):
        got = p["C"].download().reshape(M, N)
        if M == 0 or N == 0:
            continue
        rows = None if rows_per is None else np.unique(np.linspace(0, M - 1, min(M, rows_per)).astype(np.int64))
        ref, mag = ora.gemm64(A, B, M, N, K, bool(ta), bool(tb), alpha, beta, C0, rows=rows)
        g = got if rows is None else got[rows]
        cw, _, _ = ora.error_metrics(np.ascontiguousarray(g), ref, mag)
        assert cw <= TOL, ((M, N, K, ta, tb), cw)
        worst = max(worst, cw)
    return worst


def _ragged_dims(rng, n, lo, hi):
    dims = []
    for _ in range(n):
        M, N, K = (int(v) for v in rng.integers(lo, hi, 3))
        dims.append((M, N, K, int(rng.integers(0, 2)), int(rng.integers(0, 2))))
    return dims


@pytest.mark.parametrize("dtype", FORMS, ids=FORM_IDS)
def test_grouped_mixed_against_oracle(ctx, dtype):
    rng = np.random.default_rng(7)
    dims = _ragged_dims(rng, 24, 1, 300)
    dims += [(0, 5, 7, 0, 0), (5, 0, 7, 1, 0), (5, 7, 0, 0, 1), (256, 256, 256, 1, 1), (300, 260, 33, 0, 1)]
    _say("grouped mixed", dtype, len(dims))
    host, probs = _make_group(ctx, dims, 11, dtype)
    g = ctx.grouped(probs, dtype, use_graph=False)
    g.run()
    ctx.sync()
    worst = _check_oracle(dims, host, probs)
    print("[grouped] dtype %d worst componentwise err %.3e, %s" % (dtype, worst, g.info()))


@pytest.mark.parametrize("dtype", FORMS, ids=FORM_IDS)
def test_grouped_equals_single_launch(ctx, dtype):
    rng = np.random.default_rng(8)
    dims = _ragged_dims(rng, 12, 16, 520)
    _say("grouped vs single", dtype)
    host, probs = _make_group(ctx, dims, 12, dtype)
    g = ctx.grouped(probs, dtype, use_graph=True)
    g.run()
    ctx.sync()
    shapes = g.shapes()
    for (M, N, K, ta, tb), (A, B, C0, alpha, beta), p, s in zip(dims, host, probs, shapes):
        grouped = p["C"].download()
        p["C"].upload(C0)
        ctx.gemm(p["A"], p["B"], p["C"], M, N, K, bool(ta), bool(tb), alpha, beta, dtype=dtype,
                 cfg=SHAPE_CFG[dtype][s])
        ctx.sync()
        single = p["C"].download()
        assert np.array_equal(grouped.view(np.uint32), single.view(np.uint32)), (M, N, K, ta, tb, s)


def test_grouped_bf16_merged_tiles(ctx):
    """BF16 problems with aligned rows and K a multiple of 16 run pairs of
    whole 128 x 128 tiles as one 256 x 128 (M >= N) or 128 x 256 (N > M)
    tile with 32-deep direct-to-LDS stages: K = 32 (one stage), 48 and 80
    (a short first stage), 208 and 1040; all transpose pairs; sizes with an
    odd number of whole 128-wide spans and 64- and 32-wide edge strips
    beside the merged spans.  Each problem against the FP64 oracle of its
    bf16-rounded operands and bit for bit against a single launch of its
    shape's tile."""
    dims = []
    for i, K in enumerate((32, 48, 80, 208, 1040)):
        for j, (ta, tb) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
            M, N = ((400, 288), (304, 512), (384, 160), (256, 608))[(i + j) % 4]
            dims.append((M, N, K, ta, tb))
    _say("grouped bf16 merged", len(dims))
    host, probs = _make_group(ctx, dims, 23, _capi.GAT_BF16)
    g = ctx.grouped(probs, _capi.GAT_BF16, use_graph=True)
    g.run()
    ctx.sync()
    shapes = g.shapes()
    assert shapes.count(0) >= len(dims) // 2, shapes
    worst = _check_oracle(dims, host, probs)
    for (M, N, K, ta, tb), (A, B, C0, alpha, beta), p, s in zip(dims, host, probs, shapes):
        grouped = p["C"].download()
        p["C"].upload(C0)
        ctx.gemm(p["A"], p["B"], p["C"], M, N, K, bool(ta), bool(tb), alpha, beta, dtype=_capi.GAT_BF16,
                 cfg=SHAPE_CFG[_capi.GAT_BF16][s])
        ctx.sync()
        single = p["C"].download()
        assert np.array_equal(grouped.view(np.uint32), single.view(np.uint32)), (M, N, K, ta, tb, s)
    print("[grouped bf16 merged] worst componentwise err %.3e shapes %s tiles %s" % (worst, shapes, g.info()))


@pytest.mark.parametrize("dtype", FORMS, ids=FORM_IDS)
def test_grouped_edge_tiles_direct_lds(ctx, dtype):
    """Edge tiles of problems with K a multiple of 16 and aligned rows stage
    straight into LDS, rows past the matrix loaded as copies of its last
    rows and not stored: M and N multiples of 4 (BF16: 8) but not of 32,
    edges 4 to 60 wide, K = 48 (a short first stage), 64 and 272, all
    transpose pairs, general form.  Each problem against the FP64 oracle and
    bit for bit against a single launch of its shape's tile, whose edge
    tiles stage through registers."""
    mult = 8 if dtype == _capi.GAT_BF16 else 4
    sizes = [(196, 132), (72, 260), (140, 68), (324, 100)]
    dims = []
    for i, K in enumerate((64, 272) if dtype == _capi.GAT_BF16 else (48, 64, 272)):
        for j, (ta, tb) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
            M, N = sizes[(i + j) % 4]
            dims.append((M // mult * mult, N // mult * mult, K, ta, tb))
    _say("grouped edge tiles", dtype, len(dims))
    host, probs = _make_group(ctx, dims, 29, dtype)
    g = ctx.grouped(probs, dtype, use_graph=False)
    g.run()
    ctx.sync()
    shapes = g.shapes()
    worst = _check_oracle(dims, host, probs)
    for (M, N, K, ta, tb), (A, B, C0, alpha, beta), p, s in zip(dims, host, probs, shapes):
        grouped = p["C"].download()
        p["C"].upload(C0)
        ctx.gemm(p["A"], p["B"], p["C"], M, N, K, bool(ta), bool(tb), alpha, beta, dtype=dtype,
                 cfg=SHAPE_CFG[dtype][s])
        ctx.sync()
        single = p["C"].download()
        assert np.array_equal(grouped.view(np.uint32), single.view(np.uint32)), (M, N, K, ta, tb, s)
    print("[grouped edge] dtype %d worst componentwise err %.3e shapes %s" % (dtype, worst, shapes))


@pytest.mark.parametrize("dtype", FORMS, ids=FORM_IDS)
def test_grouped_plan_matches_launch(ctx, dtype):
    """gat_grouped_plan (host-only dry run) with a launch's own interior
    shapes gives the tile list that launch runs: the same number of tiles
    as gat_grouped_info reports, on real device operands."""
    rng = np.random.default_rng(0x67525550)
    dims = _workload_dims(rng, 48)
    _say("grouped plan vs launch", dtype)
    host, probs = _make_group(ctx, dims, 31, dtype)
    g = ctx.grouped(probs, dtype, use_graph=False)
    shapes = g.shapes()
    descs = ctx.gemm_descs(probs)
    sh = (ctypes.c_int * len(dims))(*shapes)
    n = ctx.lib.gat_grouped_plan(dtype, descs, len(dims), sh, None, 0)
    assert n == g.info()["tiles"], (n, g.info())
    g.close()


def test_grouped_graph_replay(ctx):
    rng = np.random.default_rng(9)
    dims = _ragged_dims(rng, 20, 8, 400)
    _say("grouped graph replay")
    host, probs = _make_group(ctx, dims, 13, general=False)  # beta = 0: replays are idempotent
    eager = ctx.grouped(probs, _capi.GAT_F32, use_graph=False)
    eager.run()
    ctx.sync()
    want = [p["C"].download() for p in probs]
    graph = ctx.grouped(probs, _capi.GAT_F32, use_graph=True)
    junk = ctx.empty(1 << 20)
    for rep in range(4):
        for p in probs:  # eager work between replays: clobber every output
            ctx.fill_uniform(p["C"], 99, rep)
        ctx.fill_uniform(junk, 98, rep)
        ctx.sync()
        graph.run()
        ctx.sync()
        for w, p in zip(want, probs):
            assert np.array_equal(p["C"].download().view(np.uint32), w.view(np.uint32)), rep
    # back-to-back launches with no host wait, outputs clobbered before each:
    # a launch that did not start from a zeroed counter would leave the
    # clobbered values, and so would every launch after it
    for i in range(3):
        for p in probs:
            ctx.fill_uniform(p["C"], 97, 2 * i)
        graph.run()
        for p in probs:
            ctx.fill_uniform(p["C"], 97, 2 * i + 1)
        eager.run()
    ctx.sync()
    for w, p in zip(want, probs):
        assert np.array_equal(p["C"].download().view(np.uint32), w.view(np.uint32))
    _check_oracle(dims, host, probs)


def _workload_dims(rng, n):
    """M, N, K log-uniform in [64, 4096], rounded to multiples of 16."""
    v = np.exp(rng.uniform(np.log(64), np.log(4096), (n, 3)))
    v = np.clip(np.round(v / 16) * 16, 64, 4096).astype(int)
    t = rng.integers(0, 2, (n, 2))
    return [(int(a), int(b), int(c), int(x), int(y)) for (a, b, c), (x, y) in zip(v, t)]


@pytest.mark.parametrize("dtype", FORMS, ids=FORM_IDS)
def test_grouped_512_workload(ctx, dtype):
    rng = np.random.default_rng(0x67525550)
    dims = _workload_dims(rng, 512)
    _say("grouped 512", dtype, "flops %.3e" % sum(2.0 * m * n * k for m, n, k, _, _ in dims))
    probs, outs = [], []
    esz = 2 if dtype == _capi.GAT_BF16 else 4
    for i, (M, N, K, ta, tb) in enumerate(dims):
        if dtype == _capi.GAT_BF16:
            fA, fB = ctx.empty(M * K), ctx.empty(K * N)
            ctx.fill_uniform(fA, 21, 2 * i)
            ctx.fill_uniform(fB, 21, 2 * i + 1)
            dA, dB = ctx.empty(M * K, np.uint16), ctx.empty(K * N, np.uint16)
            ctx.convert_bf16(fA, dA)
            ctx.convert_bf16(fB, dB)
            fA.free()
            fB.free()
        else:
            dA, dB = ctx.empty(M * K), ctx.empty(K * N)
            ctx.fill_uniform(dA, 21, 2 * i)
            ctx.fill_uniform(dB, 21, 2 * i + 1)
        dC = ctx.empty(M * N)
        probs.append(dict(transA=ta, transB=tb, M=M, N=N, K=K, alpha=1.0, beta=0.0, A=dA, B=dB, C=dC))
    assert esz in (2, 4)
    single_before = {d: ctx.tile_for(d[0], d[1], d[2], bool(d[3]), bool(d[4]), dtype) for d in dims[:32]}
    try:
        ctx.grouped_tune(probs, dtype)
        # the grouped tuner keeps a table of its own: single-problem GEMMs in
        # the same size buckets keep their (default) tiles
        for d, cfg in single_before.items():
            c = _capi.TileCfg()
            assert ctx.lib.gat_tuned_get(ctx.h, dtype, d[3], d[4], d[0], d[1], d[2], ctypes.byref(c)) == 0, d
            assert c.as_tuple() == cfg, d
        g = ctx.grouped(probs, dtype, use_graph=True)
        g.run()
        ctx.sync()
        shapes = g.shapes()
    finally:
        ctx.tuned_clear()
    got = [p["C"].download() for p in probs]
    # per-problem launch loop with the same tile choice: bit-identical
    ref = ctx.empty(max(m * n for m, n, _, _, _ in dims))
    for i, ((M, N, K, ta, tb), p, s) in enumerate(zip(dims, probs, shapes)):
        ctx.gemm(p["A"], p["B"], ref, M, N, K, bool(ta), bool(tb), 1.0, 0.0, dtype=dtype, cfg=SHAPE_CFG[dtype][s])
        ctx.sync()
        single = ref.download()[:M * N]
        assert np.array_equal(single.view(np.uint32), got[i].view(np.uint32)), (i, M, N, K, ta, tb, s)
    # a row sample of every 16th problem against the FP64 oracle
    for i in range(0, 512, 16):
        M, N, K, ta, tb = dims[i]
        A = uniform_host(M * K, 21, 2 * i)
        B = uniform_host(K * N, 21, 2 * i + 1)
        if dtype == _capi.GAT_BF16:
            A, B = bf16_bits_to_f32(_bf16_round(A)), bf16_bits_to_f32(_bf16_round(B))
        rows = np.unique(np.linspace(0, M - 1, 4).astype(np.int64))
        r64, mag = ora.gemm64(A, B, M, N, K, bool(ta), bool(tb), 1.0, 0.0, None, rows=rows)
        cw, _, _ = ora.error_metrics(np.ascontiguousarray(got[i].reshape(M, N)[rows]), r64, mag)
        assert cw <= TOL, (i, dims[i], cw)
    print("[grouped512] dtype %d %s shapes(128x128)=%d" % (dtype, g.info(), shapes.count(0)))


def test_grouped_tune_journal_replay(ctx, tmp_path):
    """The grouped tuner journals every shape trial (-3, microseconds, -4 per
    key, the reference's benchmark journal records) and, given the same
    journal again, runs no trial and installs the same shapes."""
    rng = np.random.default_rng(31)
    dims = [(int(M), int(N), int(K), int(ta), int(tb))
            for M, N, K, ta, tb in zip(rng.choice([48, 200, 640, 1100], 8), rng.choice([64, 130, 700], 8),
                                       rng.choice([96, 512], 8), rng.integers(0, 2, 8), rng.integers(0, 2, 8))]
    _say("grouped tune journal", dims)
    host, probs = _make_group(ctx, dims, 17, general=False)
    jnl = tmp_path / "grouped.jnl"
    try:
        n1 = ctx.grouped_tune(probs, _capi.GAT_F32, journal=jnl)
        g = ctx.grouped(probs, _capi.GAT_F32, use_graph=False)
        s1 = g.shapes()
        g.close()
        recs = [ln.split("\t") for ln in jnl.read_text().splitlines()]
        ctx.tuned_clear()
        n2 = ctx.grouped_tune(probs, _capi.GAT_F32, journal=jnl)
        g = ctx.grouped(probs, _capi.GAT_F32, use_graph=False)
        s2 = g.shapes()
        g.close()
    finally:
        ctx.tuned_clear()
    assert n1 == n2 >= 1
    keys = {}
    for k, v in recs:
        keys.setdefault(k, []).append(int(v))
    for k, vals in keys.items():
        assert k.startswith("groupedf32_") and k.endswith("_"), k
        assert int(k.split("_")[-2]) in (0, 1, 2, 3), k
        assert vals[0] == -3 and vals[-1] == -4 and len(vals) == 3 and vals[1] > 0, (k, vals)
    # at most four shapes per bucket, and every bucket timed at least one
    assert n1 <= len(keys) <= 4 * n1
    assert len(jnl.read_text().splitlines()) == len(recs)  # replay appended nothing
    assert s1 == s2
    assert all(0 <= s <= 3 for s in s1)
