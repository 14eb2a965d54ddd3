"""Thin Python handle over the C ABI: a device context and device arrays.

Used by the tests and the benchmark to drive the HIP kernels through the
same entry points a C/C++ caller binds (include/gatlas.h).
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import check, load

_DT = {np.dtype(np.float32): _capi.GAT_F32, np.dtype(np.float64): _capi.GAT_F64}


def dtype_code(dt):
    return _DT[np.dtype(dt)]


class DeviceArray:
    """A device allocation owned by a Context (shape/dtype kept host-side)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(ctx.lib.gat_malloc(ctx.h, self.nbytes, ctypes.byref(p)), "gat_malloc")
        self.ptr = p.value

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.size == self.size, (host.size, self.size)
        check(self.ctx.lib.gat_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data, self.nbytes), "h2d")
        return self

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(self.ctx.lib.gat_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            check(self.ctx.lib.gat_free(self.ctx.h, self.ptr), "gat_free")
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.lib.gat_free(self.ctx.h, self.ptr)
                self.ptr = None
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        self.lib = load()
        self.h = self.lib.gat_create(device)
        if not self.h:
            raise _capi.GatError("gat_create(%d): %s" % (device, _capi.last_error()))

    def close(self):
        if self.h:
            self.lib.gat_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- memory ----
    def empty(self, shape, dtype=np.float32):
        return DeviceArray(self, shape, dtype)

    def to_device(self, host):
        host = np.ascontiguousarray(host)
        return DeviceArray(self, host.shape, host.dtype).upload(host)

    def sync(self):
        check(self.lib.gat_sync(self.h), "gat_sync")

    def info(self):
        buf = ctypes.create_string_buffer(256)
        check(self.lib.gat_device_info(self.h, buf, 256), "gat_device_info")
        return buf.value.decode()

    def max_workgroup(self):
        return self.lib.gat_max_workgroup(self.h)

    # ---- timing ----
    def timer_start(self):
        check(self.lib.gat_timer_start(self.h), "timer_start")

    def timer_stop(self):
        ms = ctypes.c_float()
        check(self.lib.gat_timer_stop(self.h, ctypes.byref(ms)), "timer_stop")
        return ms.value

    # ---- synthetic inputs ----
    def fill_uniform(self, arr, seed, stream, lo=-1.0, hi=1.0):
        check(self.lib.gat_fill_uniform(self.h, dtype_code(arr.dtype), arr.ptr, arr.size, seed, stream, lo, hi),
              "fill_uniform")

    def fill_normal(self, arr, seed, stream):
        check(self.lib.gat_fill_normal(self.h, dtype_code(arr.dtype), arr.ptr, arr.size, seed, stream),
              "fill_normal")

    def scale_columns(self, arr, rows, cols, seed, stream):
        check(self.lib.gat_scale_columns(self.h, arr.ptr, rows, cols, seed, stream), "scale_columns")

    # ---- products ----
    def gemm(self, A, B, C, M, N, K, transA=False, transB=False, alpha=1.0, beta=0.0, dtype=None,
             lda=None, ldb=None, ldc=None, cfg=None):
        dt = dtype_code(C.dtype) if dtype is None else dtype
        lda = (M if transA else K) if lda is None else lda
        ldb = (K if transB else N) if ldb is None else ldb
        ldc = N if ldc is None else ldc
        if cfg is not None:
            c = _capi.TileCfg(*cfg)
            rc = self.lib.gat_gemm_cfg(self.h, dt, int(transA), int(transB), M, N, K, alpha, A.ptr, lda, B.ptr,
                                       ldb, beta, C.ptr, ldc, ctypes.byref(c))
        else:
            rc = self.lib.gat_gemm(self.h, dt, int(transA), int(transB), M, N, K, alpha, A.ptr, lda, B.ptr, ldb,
                                   beta, C.ptr, ldc)
        check(rc, "gat_gemm")

    def gemm_strided(self, A, B, C, M, N, K, batch, transA=False, transB=False, alpha=1.0, beta=0.0,
                     dtype=None):
        dt = dtype_code(C.dtype) if dtype is None else dtype
        lda = M if transA else K
        ldb = K if transB else N
        check(self.lib.gat_gemm_strided(self.h, dt, int(transA), int(transB), M, N, K, alpha, A.ptr, lda, M * K,
                                        B.ptr, ldb, K * N, beta, C.ptr, N, M * N, batch), "gat_gemm_strided")

    def gemm_compat(self, A, B, C, M, N, K, gs, bh, vl, transA=False, transB=False, gemm=False, alpha=1.0,
                    beta=1.0, packed=1):
        check(self.lib.gat_gemm_compat(self.h, dtype_code(C.dtype), int(transA), int(transB), M, N, K, int(gemm),
                                       alpha, A.ptr, B.ptr, beta, C.ptr, gs, bh, vl, packed), "gat_gemm_compat")

    def gemv(self, A, x, y, M, N, transA=False, alpha=1.0, beta=0.0, lda=None):
        lda = (M if transA else N) if lda is None else lda
        check(self.lib.gat_gemv(self.h, dtype_code(y.dtype), int(transA), M, N, alpha, A.ptr, lda, x.ptr, beta,
                                y.ptr), "gat_gemv")

    def gemv_compat(self, A, x, y, M, N, gs, bh, vl, transA=False, gemv=False, alpha=1.0, beta=1.0, packed=1):
        check(self.lib.gat_gemv_compat(self.h, dtype_code(y.dtype), int(transA), M, N, int(gemv), alpha, A.ptr,
                                       x.ptr, beta, y.ptr, gs, bh, vl, packed), "gat_gemv_compat")

    def axpy(self, alpha, x, y, z):
        check(self.lib.gat_axpy(self.h, dtype_code(z.dtype), z.size, alpha, x.ptr, y.ptr, z.ptr), "gat_axpy")

    def axpy_compat(self, alpha, x, y, z, M, N, gH, gW, bh, bw, vl, packed=1):
        check(self.lib.gat_axpy_compat(self.h, dtype_code(z.dtype), M, N, alpha, x.ptr, y.ptr, z.ptr, gH, gW, bh,
                                       bw, vl, packed), "gat_axpy_compat")

    def convert_bf16(self, src, dst):
        """FP32 device array -> BF16 (stored in a uint16 device array)."""
        check(self.lib.gat_convert_f32_bf16(self.h, src.ptr, dst.ptr, src.size), "convert_bf16")

    def convert_bf16_n(self, src, dst, n):
        """The first n elements of src -> BF16."""
        assert n <= src.size and n <= dst.size
        check(self.lib.gat_convert_f32_bf16(self.h, src.ptr, dst.ptr, n), "convert_bf16")

    def quantize_fp8(self, src, rows, cols, ld, trans, out, scales):
        check(self.lib.gat_quantize_fp8(self.h, src.ptr, rows, cols, ld, int(trans), out.ptr, scales.ptr),
              "quantize_fp8")

    def gemm_fp8(self, A8, sA, B8, sB, C, M, N, K, alpha=1.0, beta=0.0, ldc=None):
        check(self.lib.gat_gemm_fp8(self.h, M, N, K, alpha, A8.ptr, sA.ptr, B8.ptr, sB.ptr, beta, C.ptr,
                                    N if ldc is None else ldc), "gat_gemm_fp8")

    # ---- grouped GEMM ----
    @staticmethod
    def gemm_descs(problems):
        """problems: sequence of dicts with transA, transB, M, N, K, alpha, A, B,
        beta, C (DeviceArrays) and optional lda, ldb, ldc."""
        arr = (_capi.GemmDesc * max(1, len(problems)))()
        for i, p in enumerate(problems):
            ta, tb = int(p.get("transA", 0)), int(p.get("transB", 0))
            M, N, K = p["M"], p["N"], p["K"]
            arr[i] = _capi.GemmDesc(ta, tb, M, N, K, p.get("alpha", 1.0), p["A"].ptr,
                                    p.get("lda", M if ta else K), p["B"].ptr, p.get("ldb", K if tb else N),
                                    p.get("beta", 0.0), p["C"].ptr, p.get("ldc", N))
        return arr

    def gemv_grouped(self, problems):
        """problems: dicts with M, N, A, x, y (FP32 DeviceArrays) and optional
        alpha, beta, lda."""
        arr = (_capi.GemvDesc * max(1, len(problems)))()
        for i, p in enumerate(problems):
            arr[i] = _capi.GemvDesc(p["M"], p["N"], p.get("alpha", 1.0), p["A"].ptr, p.get("lda", p["N"]),
                                    p["x"].ptr, p.get("beta", 0.0), p["y"].ptr)
        check(self.lib.gat_gemv_grouped(self.h, arr, len(problems)), "gat_gemv_grouped")

    def grouped(self, problems, dtype=_capi.GAT_F32, use_graph=True):
        return Grouped(self, problems, dtype, use_graph)

    def tune_gemm(self, M, N, K, transA=False, transB=False, dtype=_capi.GAT_F32, journal="gat_tune.journal",
                  verbose=False):
        """Times every MFMA tile-table entry for one shape, then confirms the
        winner (EM fixed point), journaled; the winner is installed for the
        shape's size bucket.  Returns (cfg, GFLOP/s)."""
        c, g = _capi.TileCfg(), ctypes.c_double()
        check(self.lib.gat_tune_gemm(self.h, dtype, int(transA), int(transB), M, N, K, journal.encode(),
                                     int(verbose), ctypes.byref(c), ctypes.byref(g)), "gat_tune_gemm")
        return c.as_tuple(), g.value

    def tile_for(self, M, N, K, transA=False, transB=False, dtype=_capi.GAT_F32):
        """The tile configuration gat_gemm uses for this shape (tuned or default)."""
        c = _capi.TileCfg()
        self.lib.gat_tuned_get(self.h, dtype, int(transA), int(transB), M, N, K, ctypes.byref(c))
        return c.as_tuple()

    def tuned_clear(self):
        check(self.lib.gat_tuned_clear(self.h), "gat_tuned_clear")

    def grouped_tune(self, problems, dtype=_capi.GAT_F32, reps=2, journal=None):
        """Tunes the grouped tile shape of every size bucket the problems
        touch; `journal` (a path) records every trial and replays those
        already recorded.  Returns the number of buckets tuned."""
        descs = self.gemm_descs(problems)
        n = self.lib.gat_grouped_tune(self.h, dtype, descs, len(problems), reps,
                                      None if journal is None else str(journal).encode())
        if n < 0:
            raise _capi.GatError("grouped_tune: " + _capi.last_error())
        return n

    @staticmethod
    def tile_cfgs(dtype):
        lib = load()
        out = []
        for i in range(lib.gat_tile_cfg_count(dtype)):
            c = _capi.TileCfg()
            check(lib.gat_tile_cfg_get(dtype, i, ctypes.byref(c)), "tile_cfg_get")
            out.append(c.as_tuple())
        return out


class Grouped:
    """A persistent grouped-GEMM plan (optionally a captured HIP graph)."""

    def __init__(self, ctx, problems, dtype, use_graph):
        self.ctx = ctx
        self.n = len(problems)
        self._descs = ctx.gemm_descs(problems)
        self.h = ctx.lib.gat_grouped_create(ctx.h, dtype, self._descs, self.n, int(use_graph))
        if not self.h:
            raise _capi.GatError("gat_grouped_create: " + _capi.last_error())

    def run(self):
        check(self.ctx.lib.gat_grouped_run(self.h), "gat_grouped_run")

    def info(self):
        t, f, g = ctypes.c_size_t(), ctypes.c_double(), ctypes.c_int()
        check(self.ctx.lib.gat_grouped_info(self.h, ctypes.byref(t), ctypes.byref(f), ctypes.byref(g)), "info")
        return {"tiles": t.value, "flops": f.value, "grid": g.value}

    def shapes(self):
        return [self.ctx.lib.gat_grouped_shape(self.h, i) for i in range(self.n)]

    def close(self):
        if self.h:
            self.ctx.lib.gat_grouped_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop of the C ABI (gat_server_*): a Poisson trace of GEMM /
    GEMV requests over resident operands, admitted by HBM footprint."""

    def __init__(self, ctx, **kw):
        self.ctx = ctx
        cfg = _capi.ServeCfg()
        ctx.lib.gat_serve_cfg_default(ctypes.byref(cfg))
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise KeyError(k)
            setattr(cfg, k, v)
        self.cfg = cfg
        self.h = ctx.lib.gat_server_create(ctx.h, ctypes.byref(cfg))
        if not self.h:
            raise _capi.GatError("gat_server_create: " + _capi.last_error())

    def run(self):
        st = _capi.ServeStats()
        check(self.ctx.lib.gat_server_run(self.h, ctypes.byref(st)), "gat_server_run")
        return st.as_dict()

    def kept(self):
        out = []
        for i in range(self.ctx.lib.gat_server_kept(self.h)):
            r = _capi.ServeReq()
            check(self.ctx.lib.gat_server_kept_get(self.h, i, ctypes.byref(r)), "kept_get")
            out.append({n: getattr(r, n) for n, _ in r._fields_})
        return out

    def check_alone(self):
        m = ctypes.c_size_t()
        check(self.ctx.lib.gat_server_check_alone(self.h, ctypes.byref(m)), "check_alone")
        return m.value

    def read(self, ptr, count):
        """Copy `count` FP32 values from a device address the server owns."""
        out = np.empty(count, dtype=np.float32)
        check(self.ctx.lib.gat_memcpy_d2h(self.ctx.h, out.ctypes.data, ptr, out.nbytes), "d2h")
        return out

    def close(self):
        if self.h:
            self.ctx.lib.gat_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
