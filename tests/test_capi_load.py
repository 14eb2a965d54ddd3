"""The C ABI library loads without a GPU and exports every declared symbol."""
import os
import re

from gatlas_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gat_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    names = _declared("gatlas.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    names = _declared("gatlas.h")
    missing = [n for n in names if n not in _capi.SIGNATURES]
    assert not missing, missing


def test_no_device_reports_error_not_crash():
    # Without a GPU (this container) gat_create must fail cleanly.
    lib = _capi.load()
    if lib.gat_device_count() == 0:
        assert not lib.gat_create(0)
        assert _capi.last_error()


def test_tile_cfg_tables_are_static():
    lib = _capi.load()
    for dt in (_capi.GAT_F32, _capi.GAT_F64):
        n = lib.gat_tile_cfg_count(dt)
        assert n >= 4
        import ctypes
        c = _capi.TileCfg()
        assert lib.gat_tile_cfg_get(dt, 0, ctypes.byref(c)) == 0
        assert c.waves_m >= 1 and c.blk in (16, 32)
        assert lib.gat_tile_cfg_get(dt, n, ctypes.byref(c)) != 0


def test_ctypes_structs_match_c_layout(tmp_path):
    """Every struct of the header has the same size and field offsets in C
    (gcc) as in the ctypes mirror."""
    import ctypes
    import shutil
    import subprocess
    import pytest
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"gat_tile_cfg": _capi.TileCfg, "gat_gemm_desc": _capi.GemmDesc, "gat_gemv_desc": _capi.GemvDesc,
               "gat_serve_cfg": _capi.ServeCfg, "gat_serve_stats": _capi.ServeStats,
               "gat_serve_req": _capi.ServeReq, "gat_serve_arrival": _capi.ServeArrival}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gatlas.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        for f, _ in py._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    want = []
    for py in structs.values():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert got == want
