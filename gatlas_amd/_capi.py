"""ctypes binding of libgatlas_hip.so (declarations in include/gatlas.h).

Loading fails loudly when the shared library is missing: there is no CPU
fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GATLAS_LIB names another build of the same library (e.g. the host-sanitised
# one the CPU tests run against); by default the in-tree build
LIB_PATH = os.environ.get("GATLAS_LIB") or os.path.join(_HERE, "libgatlas_hip.so")

GAT_F32, GAT_F64, GAT_BF16, GAT_BF16X3, GAT_FP8 = 0, 1, 2, 3, 4
GAT_MATMUL, GAT_MATVEC, GAT_SAXPY = 0, 1, 2
GAT_BUFFER, GAT_IMAGE, GAT_MFMA, GAT_NATIVE = 0, 1, 2, 3
GAT_OK, GAT_ERR_INVALID, GAT_ERR_HIP, GAT_ERR_NOMEM, GAT_ERR_UNSUPPORTED = 0, -1, -2, -3, -4

_c_sz = ctypes.c_size_t
_c_vp = ctypes.c_void_p
_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_c_u64 = ctypes.c_uint64


class TileCfg(ctypes.Structure):
    _fields_ = [("waves_m", _c_int), ("waves_n", _c_int), ("tm", _c_int), ("tn", _c_int),
                ("bk", _c_int), ("blk", _c_int), ("wk", _c_int)]

    def as_tuple(self):
        """(waves_m, waves_n, tm, tn, bk, blk, wk), wk normalised to >= 1."""
        return (self.waves_m, self.waves_n, self.tm, self.tn, self.bk, self.blk, max(1, self.wk))


class GemmDesc(ctypes.Structure):
    _fields_ = [("transA", _c_int), ("transB", _c_int), ("M", _c_sz), ("N", _c_sz), ("K", _c_sz),
                ("alpha", _c_dbl), ("A", _c_vp), ("lda", _c_sz), ("B", _c_vp), ("ldb", _c_sz),
                ("beta", _c_dbl), ("C", _c_vp), ("ldc", _c_sz)]


class GemvDesc(ctypes.Structure):
    _fields_ = [("M", _c_sz), ("N", _c_sz), ("alpha", _c_dbl), ("A", _c_vp), ("lda", _c_sz), ("x", _c_vp),
                ("beta", _c_dbl), ("y", _c_vp)]


GAT_GEMV_NOALLOC_SCRATCH = 16 << 20   # bytes of static scratch per GPU (gatlas.h)
GAT_SERVE_BATCHED = 0
GAT_SERVE_LOOP = 1


class ServeCfg(ctypes.Structure):
    _fields_ = [("rate", _c_dbl), ("duration", _c_dbl), ("gemm_fraction", _c_dbl), ("hbm_budget", _c_sz),
                ("mode", _c_int), ("streams", _c_int), ("max_batch", _c_int), ("max_inflight", _c_int),
                ("gemm_min", _c_int), ("gemm_max", _c_int), ("gemv_min", _c_int), ("gemv_max", _c_int),
                ("keep", _c_int), ("keep_max", _c_int), ("seed", ctypes.c_ulonglong),
                ("gemm_dtype", _c_int)]


class ServeStats(ctypes.Structure):
    _fields_ = [(n, _c_sz) for n in ("requests", "gemm_requests", "gemv_requests", "launches", "kept")] + \
               [(n, _c_dbl) for n in ("wall_s", "flops", "bytes", "tflops", "req_per_s", "p50_ms", "p99_ms",
                                      "max_ms", "p50_gemm_ms", "p99_gemm_ms", "p50_gemv_ms", "p99_gemv_ms",
                                      "mean_queue_ms", "mean_batch", "peak_footprint", "device_busy", "blocked_ms",
                                      "offered_tflops", "peak_output_bytes")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServeArrival(ctypes.Structure):
    _fields_ = [("t", _c_dbl), ("kind", _c_int), ("transA", _c_int), ("transB", _c_int), ("M", _c_sz),
                ("N", _c_sz), ("K", _c_sz), ("footprint", _c_sz)]


class ServeReq(ctypes.Structure):
    _fields_ = [("kind", _c_int), ("transA", _c_int), ("transB", _c_int), ("shape", _c_int), ("M", _c_sz),
                ("N", _c_sz), ("K", _c_sz), ("A", _c_vp), ("lda", _c_sz), ("B", _c_vp), ("ldb", _c_sz),
                ("C", _c_vp), ("ldc", _c_sz), ("latency_ms", _c_dbl), ("id", _c_sz)]


# name -> (restype, argtypes)
SIGNATURES = {
    "gat_gemv_grouped": (_c_int, [_c_vp, ctypes.POINTER(GemvDesc), _c_sz]),
    "gat_serve_cfg_default": (None, [ctypes.POINTER(ServeCfg)]),
    "gat_serve_trace": (ctypes.c_long, [ctypes.POINTER(ServeCfg), ctypes.POINTER(ServeArrival), _c_sz]),
    "gat_server_create": (_c_vp, [_c_vp, ctypes.POINTER(ServeCfg)]),
    "gat_server_run": (_c_int, [_c_vp, ctypes.POINTER(ServeStats)]),
    "gat_server_kept": (_c_sz, [_c_vp]),
    "gat_server_kept_get": (_c_int, [_c_vp, _c_sz, ctypes.POINTER(ServeReq)]),
    "gat_server_check_alone": (_c_int, [_c_vp, ctypes.POINTER(_c_sz)]),
    "gat_server_destroy": (None, [_c_vp]),
    "gat_grouped_create": (_c_vp, [_c_vp, _c_int, ctypes.POINTER(GemmDesc), _c_sz, _c_int]),
    "gat_grouped_run": (_c_int, [_c_vp]),
    "gat_grouped_info": (_c_int, [_c_vp, ctypes.POINTER(_c_sz), ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_int)]),
    "gat_grouped_shape": (_c_int, [_c_vp, _c_sz]),
    "gat_grouped_plan": (ctypes.c_long, [_c_int, ctypes.POINTER(GemmDesc), _c_sz, ctypes.POINTER(_c_int),
                                         ctypes.POINTER(_c_int), _c_sz]),
    "gat_grouped_destroy": (None, [_c_vp]),
    "gat_gemm_grouped": (_c_int, [_c_vp, _c_int, ctypes.POINTER(GemmDesc), _c_sz]),
    "gat_grouped_tune": (_c_int, [_c_vp, _c_int, ctypes.POINTER(GemmDesc), _c_sz, _c_int, ctypes.c_char_p]),
    "gat_create": (_c_vp, [_c_int]),
    "gat_destroy": (None, [_c_vp]),
    "gat_last_error": (ctypes.c_char_p, []),
    "gat_max_workgroup": (_c_sz, [_c_vp]),
    "gat_device_info": (_c_int, [_c_vp, ctypes.c_char_p, _c_sz]),
    "gat_stream": (_c_vp, [_c_vp]),
    "gat_device_count": (_c_int, []),
    "gat_malloc": (_c_int, [_c_vp, _c_sz, ctypes.POINTER(_c_vp)]),
    "gat_free": (_c_int, [_c_vp, _c_vp]),
    "gat_memcpy_h2d": (_c_int, [_c_vp, _c_vp, _c_vp, _c_sz]),
    "gat_memcpy_d2h": (_c_int, [_c_vp, _c_vp, _c_vp, _c_sz]),
    "gat_memcpy_d2d": (_c_int, [_c_vp, _c_vp, _c_vp, _c_sz]),
    "gat_memset": (_c_int, [_c_vp, _c_vp, _c_int, _c_sz]),
    "gat_sync": (_c_int, [_c_vp]),
    "gat_timer_start": (_c_int, [_c_vp]),
    "gat_timer_stop": (_c_int, [_c_vp, ctypes.POINTER(ctypes.c_float)]),
    "gat_fill_uniform": (_c_int, [_c_vp, _c_int, _c_vp, _c_sz, _c_u64, _c_u64, _c_dbl, _c_dbl]),
    "gat_fill_normal": (_c_int, [_c_vp, _c_int, _c_vp, _c_sz, _c_u64, _c_u64]),
    "gat_scale_columns": (_c_int, [_c_vp, _c_vp, _c_sz, _c_sz, _c_u64, _c_u64]),
    "gat_gemm": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, _c_dbl, _c_vp, _c_sz, _c_vp,
                          _c_sz, _c_dbl, _c_vp, _c_sz]),
    "gat_gemm_strided": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, _c_dbl, _c_vp, _c_sz,
                                  _c_sz, _c_vp, _c_sz, _c_sz, _c_dbl, _c_vp, _c_sz, _c_sz, _c_sz]),
    "gat_gemm_cfg": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, _c_dbl, _c_vp, _c_sz,
                              _c_vp, _c_sz, _c_dbl, _c_vp, _c_sz, ctypes.POINTER(TileCfg)]),
    "gat_tile_cfg_count": (_c_int, [_c_int]),
    "gat_tile_cfg_get": (_c_int, [_c_int, _c_int, ctypes.POINTER(TileCfg)]),
    "gat_gemm_compat": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, _c_int, _c_dbl, _c_vp,
                                 _c_vp, _c_dbl, _c_vp, _c_sz, _c_sz, _c_sz, _c_sz]),
    "gat_gemv": (_c_int, [_c_vp, _c_int, _c_int, _c_sz, _c_sz, _c_dbl, _c_vp, _c_sz, _c_vp, _c_dbl, _c_vp]),
    "gat_gemv_noalloc": (_c_int, [_c_vp, _c_int, _c_int, _c_sz, _c_sz, _c_dbl, _c_vp, _c_sz, _c_vp, _c_dbl,
                                  _c_vp]),
    "gat_gemv_compat": (_c_int, [_c_vp, _c_int, _c_int, _c_sz, _c_sz, _c_int, _c_dbl, _c_vp, _c_vp, _c_dbl,
                                 _c_vp, _c_sz, _c_sz, _c_sz, _c_sz]),
    "gat_axpy": (_c_int, [_c_vp, _c_int, _c_sz, _c_dbl, _c_vp, _c_vp, _c_vp]),
    "gat_axpy_compat": (_c_int, [_c_vp, _c_int, _c_sz, _c_sz, _c_dbl, _c_vp, _c_vp, _c_vp, _c_sz, _c_sz,
                                 _c_sz, _c_sz, _c_sz, _c_sz]),
    "gat_convert_f32_bf16": (_c_int, [_c_vp, _c_vp, _c_vp, _c_sz]),
    "gat_quantize_fp8": (_c_int, [_c_vp, _c_vp, _c_sz, _c_sz, _c_sz, _c_int, _c_vp, _c_vp]),
    "gat_gemm_fp8": (_c_int, [_c_vp, _c_sz, _c_sz, _c_sz, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp, _c_sz]),
    # parameter interface
    "gat_kernel_new": (_c_vp, [_c_vp, _c_int, _c_int, _c_int, _c_int]),
    "gat_kernel_free": (None, [_c_vp]),
    "gat_kernel_name": (ctypes.c_char_p, [_c_vp]),
    "gat_kernel_param_count": (_c_int, [_c_vp]),
    "gat_kernel_set_params": (_c_int, [_c_vp, ctypes.POINTER(_c_sz), _c_sz]),
    "gat_kernel_valid": (_c_int, [_c_vp]),
    "gat_kernel_get_params": (_c_int, [_c_vp, ctypes.POINTER(_c_sz), _c_sz]),
    "gat_kernel_set_paranoid": (_c_int, [_c_vp]),
    "gat_kernel_set_packed": (_c_int, [_c_vp, _c_sz]),
    "gat_kernel_total_variations": (_c_sz, [_c_vp]),
    "gat_kernel_flops": (_c_sz, [_c_vp]),
    "gat_kernel_extra_detail": (_c_int, [_c_vp, ctypes.POINTER(_c_sz), _c_sz]),
    "gat_kernel_work_items": (_c_int, [_c_vp, ctypes.POINTER(_c_sz), ctypes.POINTER(_c_sz), ctypes.POINTER(_c_sz),
                                       ctypes.POINTER(_c_sz), _c_sz]),
    "gat_kernel_print": (ctypes.c_char_p, [_c_vp]),
    "gat_kernel_build": (_c_int, [_c_vp]),
    "gat_kernel_set_args": (_c_int, [_c_vp, _c_int]),
    "gat_kernel_run": (_c_int, [_c_vp, _c_sz, ctypes.POINTER(_c_dbl)]),
    "gat_kernel_sync_output": (_c_int, [_c_vp]),
    "gat_kernel_check": (_c_int, [_c_vp, _c_int, ctypes.POINTER(_c_dbl)]),
    # journal
    "gat_journal_open": (_c_vp, [ctypes.c_char_p]),
    "gat_journal_close": (None, [_c_vp]),
    "gat_journal_load": (_c_int, [_c_vp]),
    "gat_journal_purge": (_c_int, [_c_vp, _c_int]),
    "gat_journal_good": (_c_sz, [_c_vp]),
    "gat_journal_state": (_c_int, [_c_vp, _c_vp, ctypes.POINTER(_c_sz), _c_sz]),
    "gat_journal_time": (_c_int, [_c_vp, _c_vp, ctypes.POINTER(_c_sz), _c_sz, _c_sz]),
    "gat_journal_take": (_c_int, [_c_vp, _c_vp, ctypes.POINTER(_c_sz), _c_sz, _c_int]),
    # ranking / search / tuning
    "gat_rank_bench": (_c_int, [_c_sz, ctypes.POINTER(_c_int), ctypes.POINTER(_c_dbl), _c_sz]),
    "gat_mark_bench": (_c_int, [_c_sz, ctypes.POINTER(_c_int), ctypes.POINTER(_c_sz), _c_sz]),
    "gat_mark_bench_avg": (_c_int, [_c_sz, ctypes.POINTER(_c_int), ctypes.POINTER(_c_dbl), _c_sz]),
    "gat_search_params": (_c_int, [_c_vp, _c_sz, _c_sz, _c_sz, _c_sz, _c_int, _c_sz, _c_sz, _c_sz, _c_int, _c_int,
                                   ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.POINTER(_c_sz), _c_sz]),
    "gat_search_em": (_c_int, [_c_vp, ctypes.c_char_p, _c_sz, _c_sz, _c_sz, _c_sz, _c_int, _c_sz, _c_sz, _c_sz,
                               _c_int, _c_int, _c_sz, _c_int, ctypes.POINTER(_c_sz), _c_sz]),
    "gat_tune_gemm": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, ctypes.c_char_p, _c_int,
                               ctypes.POINTER(TileCfg), ctypes.POINTER(_c_dbl)]),
    "gat_tuned_set": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, ctypes.POINTER(TileCfg)]),
    "gat_tuned_get": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_sz, _c_sz, _c_sz, ctypes.POINTER(TileCfg)]),
    "gat_tuned_clear": (_c_int, [_c_vp]),
}

_lib = None


class GatError(RuntimeError):
    pass


def load():
    """Load libgatlas_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libgatlas_hip.so is not built (expected at %s); run __graft_entry__.build() or "
            "make -C gatlas_amd/csrc" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    msg = load().gat_last_error()
    return msg.decode() if msg else ""


def check(rc, what=""):
    if rc != 0:
        raise GatError("%s failed (%d): %s" % (what, rc, last_error()))
    return rc
