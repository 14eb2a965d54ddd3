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

    @staticmethod
    def tile_cfgs(dtype):
        lib = load()
        out = []
        for i in range(lib.gat_tile_cfg_count(dtype)):
            c = _capi.TileCfg()
            check(lib.gat_tile_cfg_get(dtype, i, ctypes.byref(c)), "tile_cfg_get")
            out.append(c.as_tuple())
        return out
