"""The serving trace (gat_serve_trace) on the CPU: Poisson arrivals, request
mix, shape distributions and footprints as the served workload specifies
(arrivals at `rate` per second; 70% GEMM with M, N, K log-uniform in
[64, 4096] rounded to multiples of 16 and random transposes; 30% GEMV with
M, N log-uniform in [1024, 32768], multiples of 16)."""
import ctypes

import numpy as np
import pytest
from scipy import stats

from gatlas_amd import _capi


def _trace(**kw):
    lib = _capi.load()
    cfg = _capi.ServeCfg()
    lib.gat_serve_cfg_default(ctypes.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    n = lib.gat_serve_trace(ctypes.byref(cfg), None, 0)
    assert n >= 0, _capi.last_error()
    arr = (_capi.ServeArrival * max(1, n))()
    assert lib.gat_serve_trace(ctypes.byref(cfg), arr, n) == n
    return cfg, [{f: getattr(a, f) for f, _ in a._fields_} for a in arr[:n]]


def test_poisson_arrivals_and_mix():
    rate, dur = 10000.0, 2.0
    cfg, tr = _trace(rate=rate, duration=dur, seed=11)
    n = len(tr)
    assert abs(n - rate * dur) < 5 * np.sqrt(rate * dur)
    t = np.array([r["t"] for r in tr])
    assert np.all(np.diff(t) > 0) and t[0] >= 0 and t[-1] < dur
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert stats.kstest(gaps * rate, "expon").pvalue > 1e-4
    g = np.array([r["kind"] == 0 for r in tr])
    assert abs(g.mean() - 0.7) < 5 * np.sqrt(0.21 / n)


def test_shapes_and_footprints():
    cfg, tr = _trace(rate=20000.0, duration=1.0, seed=12)
    gemm = [r for r in tr if r["kind"] == 0]
    gemv = [r for r in tr if r["kind"] == 1]
    for r in gemm:
        for d in (r["M"], r["N"], r["K"]):
            assert 64 <= d <= 4096 and d % 16 == 0
        assert r["footprint"] == 4 * (r["M"] * r["K"] + r["K"] * r["N"] + r["M"] * r["N"])
    for r in gemv:
        assert 1024 <= r["M"] <= 32768 and 1024 <= r["N"] <= 32768
        assert r["M"] % 16 == 0 and r["N"] % 16 == 0 and r["K"] == 0
        assert r["footprint"] == 4 * (r["M"] * r["N"] + r["N"] + r["M"])
    # log-uniform: log of the (unrounded-equivalent) sizes is close to uniform
    lm = np.log(np.array([r["K"] for r in gemm], dtype=float))
    u = (lm - np.log(64)) / (np.log(4096) - np.log(64))
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    lv = np.log(np.array([r["M"] for r in gemv], dtype=float))
    u = (lv - np.log(1024)) / (np.log(32768) - np.log(1024))
    assert abs(u.mean() - 0.5) < 0.03
    ta = np.array([r["transA"] for r in gemm])
    tb = np.array([r["transB"] for r in gemm])
    assert abs(ta.mean() - 0.5) < 0.03 and abs(tb.mean() - 0.5) < 0.03


def test_trace_is_deterministic_in_the_seed():
    _, a = _trace(rate=2000.0, duration=0.5, seed=5)
    _, b = _trace(rate=2000.0, duration=0.5, seed=5)
    _, c = _trace(rate=2000.0, duration=0.5, seed=6)
    assert a == b and a != c


@pytest.mark.parametrize("bad", [dict(rate=0.0), dict(duration=-1.0), dict(gemm_fraction=1.5), dict(streams=0),
                                 dict(streams=17), dict(gemm_min=8), dict(gemv_max=100000), dict(mode=7),
                                 dict(rate=1e7, duration=1.0)])
def test_invalid_configuration_is_rejected(bad):
    lib = _capi.load()
    cfg = _capi.ServeCfg()
    lib.gat_serve_cfg_default(ctypes.byref(cfg))
    for k, v in bad.items():
        setattr(cfg, k, v)
    assert lib.gat_serve_trace(ctypes.byref(cfg), None, 0) < 0
    assert _capi.last_error()
