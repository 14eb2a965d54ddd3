"""getParams candidate sets at the benchmark workloads' sizes, pinned by brute
force: every point of the whole (work group, block height, extra parameter)
grid is tested with validParams -- the predicate pinned against the compiled
reference's own vectors (tests/golden/ref_params.jsonl) -- and the set the
reference's enumeration keeps is formed from the valid points declaratively:

  * GEMM (benchMatmul.cpp:258-298): the valid points of the largest square
    group above 8 (at most min(maxGroupSize, sqrt(maxWorkGroupSize))) that has
    any, plus those of the 8 x 8 group;
  * GEMV (benchMatvec.cpp:243-283): the same with groups above 64 (at most
    min(maxGroupSize, maxWorkGroupSize)) plus the 64 group;
  * SAXPY, 1-D (benchSaxpy.cpp:232-246): every valid point with a group of
    64 .. min(maxGroupSize, maxWorkGroupSize).
The enumerator's list must hold exactly that set, once each, in the
enumeration's nesting order (group, then block height, then extra
parameter).  The bandwidth kernels' own launch-shape variant (NATIVE) has
no reference counterpart and is not covered by these rules.  Runs on the
CPU: the parameter functions need no device."""
import ctypes

import pytest

from gatlas_amd import _capi
from gatlas_amd.interface import BUFFER, IMAGE, MATMUL, MATVEC, SAXPY, Kernel

MAXWG = 1024   # MI355X maxThreadsPerBlock (the reference's maxWorkGroupSize)


def _valid(k, p):
    arr = (ctypes.c_size_t * len(p))(*p)
    return k.lib.gat_kernel_set_params(k.h, arr, len(p)) == 0 and k.lib.gat_kernel_valid(k.h) == 1


def _order_key_matmul(p):
    return (-p[7] if p[7] != 8 else 0, p[9], p[11])   # largest group first, the 8 x 8 group last


@pytest.mark.parametrize("variant,vl,maxbh,maxgs", [(BUFFER, 4, 10, 10), (BUFFER, 2, 10, 10),
                                                    (IMAGE, 4, 12, 16)])
@pytest.mark.parametrize("n", [1024, 2048, 4096, 8192])
@pytest.mark.parametrize("gemm", [0, 1])
def test_matmul_candidate_set(variant, vl, maxbh, maxgs, n, gemm):
    k = Kernel(MATMUL, _capi.GAT_F32, vl, variant)
    nxp = k.totalVariations()
    largest = min(maxgs, int(MAXWG ** 0.5))
    for ta in (0, 1):
        for tb in (0, 1):
            grid = {}
            for wg in range(8, largest + 1):
                grid[wg] = [p for bh in range(vl, maxbh + 1) for xp in range(nxp)
                            for p in [[1, gemm, n, n, n, ta, tb, wg, wg, bh, vl, xp]] if _valid(k, p)]
            top = max([wg for wg in range(9, largest + 1) if grid[wg]], default=None)
            want = (grid[top] if top else []) + grid[8]
            got = k.search(n, n, n, bool(ta), bool(tb), bool(gemm), vl=vl, max_workgroup=MAXWG,
                           max_block_height=maxbh, max_group_size=maxgs)
            assert sorted(map(tuple, got)) == sorted(map(tuple, want)), (n, ta, tb)
            assert len(set(map(tuple, got))) == len(got)
            assert got == sorted(got, key=_order_key_matmul)
            assert len(got) > 0


@pytest.mark.parametrize("variant,vl", [(BUFFER, 4), (BUFFER, 1), (IMAGE, 4)])
@pytest.mark.parametrize("ta", [0, 1])
def test_matvec_candidate_set_16384(variant, vl, ta):
    m = 16384
    k = Kernel(MATVEC, _capi.GAT_F32, vl, variant)
    nxp = k.totalVariations()
    maxbh, maxgs = 64, 256
    largest = min(maxgs, MAXWG)
    grid = {}
    for wg in range(64, largest + 1):
        grid[wg] = [p for bh in range(vl, maxbh + 1) for xp in range(nxp)
                    for p in [[1, 1, m, m, ta, wg, bh, vl, xp]] if _valid(k, p)]
    top = max([wg for wg in range(65, largest + 1) if grid[wg]], default=None)
    want = (grid[top] if top else []) + grid[64]
    got = k.search(m, m, 0, bool(ta), False, True, vl=vl, max_workgroup=MAXWG, max_block_height=maxbh,
                   max_group_size=maxgs)
    assert sorted(map(tuple, got)) == sorted(map(tuple, want))
    assert len(set(map(tuple, got))) == len(got) > 0
    assert got == sorted(got, key=lambda p: (-p[5] if p[5] != 64 else 0, p[6], p[8]))


@pytest.mark.parametrize("variant,vl", [(BUFFER, 4), (BUFFER, 1), (IMAGE, 4)])
def test_saxpy_candidate_set_2_28(variant, vl):
    n = 1 << 28
    M, N = n // vl, vl            # 1-D form: N equals the vector length
    k = Kernel(SAXPY, _capi.GAT_F32, vl, variant)
    nxp = k.totalVariations()
    maxbh, maxgs = 16, 256
    largest = min(maxgs, MAXWG)
    want = [p for wg in range(64, largest + 1) for bh in range(1, maxbh + 1) for xp in range(nxp)
            for p in [[1, M, N, wg, 0, bh, 0, vl, xp]] if _valid(k, p)]
    got = k.search(M, N, 0, False, False, False, vl=vl, max_workgroup=MAXWG, max_block_height=maxbh,
                   max_group_size=maxgs, max_out=40000)
    assert sorted(map(tuple, got)) == sorted(map(tuple, want))
    assert len(set(map(tuple, got))) == len(got) > 0
    assert got == sorted(got, key=lambda p: (p[3], p[5], p[8]))
