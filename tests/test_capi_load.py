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
