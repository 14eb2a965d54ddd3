"""The parameter interface, journal and ranking helpers of libgatlas_hip.so
against golden fixtures written by the reference's own code
(tests/golden/make_ref_golden.py, oracle/_ref).  CPU only: no device calls."""
import json
import os

import pytest

from gatlas_amd import _capi
from gatlas_amd.interface import (BUFFER, IMAGE, MATMUL, MATVEC, SAXPY, Journal, Kernel, mark_bench,
                                  mark_bench_avg, rank_bench)

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _lines(name):
    with open(os.path.join(GOLD, name)) as f:
        return [json.loads(x) for x in f]


PARAMS = _lines("ref_params.jsonl")
FAM = {"matmul": MATMUL, "matvec": MATVEC, "saxpy": SAXPY}


def _kernel(fam, variant, params):
    if variant == "image":
        vl = params[10] if fam == "matmul" else 4
        dt = _capi.GAT_F32 if vl == 4 else _capi.GAT_F64
        return Kernel(FAM[fam], dt, vl if fam == "matmul" else 4, IMAGE)
    vl = params[10] if fam == "matmul" else params[7]
    return Kernel(FAM[fam], _capi.GAT_F32, vl, BUFFER)


def test_fixture_counts():
    assert len(PARAMS) >= 2000
    assert sum(c["valid"] for c in PARAMS) >= 100


@pytest.mark.parametrize("fam", ["matmul", "matvec", "saxpy"])
def test_parameter_functions_match_reference(fam):
    n = 0
    for c in PARAMS:
        if c["family"] != fam:
            continue
        k = _kernel(fam, c["variant"], c["in"])
        k.setParams(c["in"])
        assert k.validParams() == bool(c["valid"]), c
        if c["valid"]:
            assert k.getParams() == c["params"], c
        g, l = k.workItems()
        assert g == c["global"], c
        assert l == c["local"], c
        assert k.numberFlops() == c["flops"], c
        assert k.extraParamDetail() == c["extra"], c
        assert k.totalVariations() == c["total"], c
        n += 1
    assert n > 500


def _kernel_for_name(name):
    table = {"matmulbufferfloat4": (MATMUL, 4), "matmulbufferfloat2": (MATMUL, 2),
             "matvecbufferfloat4": (MATVEC, 4), "saxpybufferfloat1": (SAXPY, 1)}
    fam, vl = table[name]
    k = Kernel(fam, _capi.GAT_F32, vl, BUFFER)
    assert k.kernelName() == name
    return k


@pytest.mark.parametrize("mode", ["purge0", "purge1"])
def test_journal_matches_reference(tmp_path, mode):
    fx = json.load(open(os.path.join(GOLD, "ref_journal.json")))[mode]
    path = str(tmp_path / "journal.txt")
    j = None
    outs = []
    for line in fx["script"]:
        tok = line.split()
        cmd = tok[0]
        if cmd == "jfile":
            j = Journal(path)
            outs.append({"ok": 1})
        elif cmd in ("jtake", "jstate", "jtime"):
            k = _kernel_for_name(tok[1])
            n = int(tok[2])
            p = [int(x) for x in tok[3:3 + n]]
            v = int(tok[3 + n]) if len(tok) > 3 + n else 0
            if cmd == "jtake":
                outs.append({"ok": int(j.takeMemo(k, p, v))})
            elif cmd == "jstate":
                outs.append({"state": j.memoRunState(k, p)})
            else:
                outs.append({"time": j.memoTime(k, p, v)})
        elif cmd == "jload":
            outs.append({"ok": int(j.loadMemo())})
        elif cmd == "jgood":
            outs.append({"good": j.memoGood()})
    assert outs == fx["outputs"]
    assert open(path).read() == fx["file_before"]
    j2 = Journal(path)
    assert j2.loadMemo()
    bad = j2.purgeMemo(mode == "purge1")
    assert {"bad": bad} == fx["purge_result"]
    del j2
    assert open(path).read() == fx["file_after"]


def test_rank_and_mark_match_reference():
    for rec in _lines("ref_rank.jsonl"):
        tok = rec["cmd"].split()
        which, a, n = tok[0], int(tok[1]), int(tok[2])
        ok = [int(x) for x in tok[3:3 + n]]
        vals = tok[3 + n:3 + 2 * n]
        if which == "rank":
            assert {"rank": rank_bench(a, ok, [float(v) for v in vals])} == rec["out"], rec
        elif which == "mark":
            assert {"ok": mark_bench(a, ok, [int(v) for v in vals])} == rec["out"], rec
        else:
            assert {"ok": mark_bench_avg(a, ok, [float(v) for v in vals])} == rec["out"], rec


def test_extra_parameter_digits_hand_cases():
    # matmulbuffer: digits (inlineMNK radix 2, loopOrder radix 6), least
    # significant first: 11 = 1 + 2*5.
    k = Kernel(MATMUL, _capi.GAT_F32, 4, BUFFER)
    k.setParams([1, 0, 64, 64, 64, 0, 0, 8, 8, 4, 4, 11])
    assert k.extraParamDetail() == [1, 5]
    assert k.totalVariations() == 12
    k.setParams([1, 0, 64, 64, 64, 0, 0, 8, 8, 4, 4, 12])
    assert not k.validParams()
    img = Kernel(MATMUL, _capi.GAT_F32, 4, IMAGE)
    img.setParams([1, 0, 64, 64, 64, 0, 0, 8, 8, 4, 4, 23])
    assert img.extraParamDetail() == [1, 5, 1] and img.totalVariations() == 24
