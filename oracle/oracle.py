"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
bench's CPU-baseline leg, always as the checker.  The product package never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t


def build():
    """Compile liboracle.so in place (gcc, no GPU needed)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_lcg48_next.argtypes = [_u64p]
        L.ora_lcg48_next.restype = ctypes.c_double
        L.ora_posrand_f32.argtypes = [_u64p]
        L.ora_posrand_f32.restype = ctypes.c_float
        L.ora_posrand_f64.argtypes = [_u64p]
        L.ora_posrand_f64.restype = ctypes.c_double
        L.ora_fillrand_f32.argtypes = [_u64p, _f32p, _sz]
        L.ora_fillrand_f64.argtypes = [_u64p, _f64p, _sz]
        for name, fp, T in (("f32", _f32p, ctypes.c_float), ("f64", _f64p, ctypes.c_double)):
            fn = getattr(L, "ora_ref_matmul_" + name)
            fn.argtypes = [ctypes.c_int] * 4 + [_sz] * 3 + [T, T, fp, fp, ctypes.c_void_p, fp]
            fn = getattr(L, "ora_ref_matvec_" + name)
            fn.argtypes = [ctypes.c_int] * 3 + [_sz] * 2 + [T, T, fp, fp, ctypes.c_void_p, fp]
            fn = getattr(L, "ora_ref_saxpy_" + name)
            fn.argtypes = [_sz, T, fp, fp, fp]
            fn = getattr(L, "ora_absdiff_" + name)
            fn.argtypes = [fp, fp, _sz]
            fn.restype = ctypes.c_longdouble
            fn = getattr(L, "ora_paranoid_pass_" + name)
            fn.argtypes = [fp, fp, _sz]
            fn.restype = ctypes.c_int
            fn = getattr(L, "ora_check_array_" + name)
            fn.argtypes = [fp, _sz, T]
            fn.restype = ctypes.c_int
            fn = getattr(L, "ora_error_metrics_" + name)
            fn.argtypes = [fp, _f64p, ctypes.c_void_p, _sz, _f64p]
        L.ora_gemm_f64.argtypes = ([ctypes.c_int] * 2 + [_sz] * 3 + [ctypes.c_double, _f64p, _sz, _f64p, _sz,
                                   ctypes.c_double, ctypes.c_void_p, _sz, ctypes.c_void_p, _sz,
                                   _f64p, ctypes.c_void_p])
        L.ora_gemm_f64_from_f32.argtypes = ([ctypes.c_int] * 2 + [_sz] * 3 + [ctypes.c_double, _f32p, _sz, _f32p, _sz,
                                            ctypes.c_double, ctypes.c_void_p, _sz, ctypes.c_void_p, _sz,
                                            _f64p, ctypes.c_void_p])
        L.ora_gemv_f64_from_f32.argtypes = [ctypes.c_int, _sz, _sz, ctypes.c_double, _f32p, _sz, _f32p,
                                            ctypes.c_double, ctypes.c_void_p, _f64p, ctypes.c_void_p]
        L.ora_port_matmul_f32_rows.argtypes = ([ctypes.c_int] * 3 + [_sz] * 3 + [ctypes.c_float] * 2 +
                                               [_f32p, _f32p, _f32p, _sz, _sz])
        L.ora_port_matvec_f32_rows.argtypes = [ctypes.c_int, _sz, ctypes.c_float, ctypes.c_float,
                                               _f32p, _f32p, _f32p, _sz, _sz]
        L.ora_port_saxpy_f32.argtypes = [_sz, _sz, ctypes.c_float, _f32p, _f32p, _f32p]
        L.ora_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Drand48:
    """The reference's posrand()/fillrand() stream (OCLAppUtil.hpp:28-39).

    The default state 0 is glibc's unseeded drand48 state."""

    def __init__(self, state=0):
        self.state = ctypes.c_uint64(state)

    def next(self):
        return lib().ora_lcg48_next(ctypes.byref(self.state))

    def posrand(self, dtype):
        if np.dtype(dtype) == np.float32:
            return np.float32(lib().ora_posrand_f32(ctypes.byref(self.state)))
        return np.float64(lib().ora_posrand_f64(ctypes.byref(self.state)))

    def fillrand(self, n, dtype):
        out = np.empty(n, dtype=dtype)
        if out.dtype == np.float32:
            lib().ora_fillrand_f32(ctypes.byref(self.state), out, n)
        else:
            lib().ora_fillrand_f64(ctypes.byref(self.state), out, n)
        return out


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def ref_matmul(A, B, M, N, K, transA=False, transB=False, gemm=False, alpha=1.0, beta=1.0,
               Cold=None, packed=1):
    """The reference's paranoid host product, bit-for-bit (scalar precision)."""
    dt = A.dtype
    C = np.zeros(packed * M * N, dtype=dt)
    fn = getattr(lib(), "ora_ref_matmul_" + _sfx(dt))
    fn(packed, int(gemm), int(transA), int(transB), M, N, K, dt.type(alpha), dt.type(beta),
       np.ascontiguousarray(A.ravel()), np.ascontiguousarray(B.ravel()),
       _ptr(None if Cold is None else np.ascontiguousarray(Cold.ravel())), C)
    return C


def ref_matvec(A, x, M, N, transA=False, gemv=False, alpha=1.0, beta=1.0, yold=None, packed=1):
    dt = A.dtype
    y = np.zeros(packed * M, dtype=dt)
    fn = getattr(lib(), "ora_ref_matvec_" + _sfx(dt))
    fn(packed, int(gemv), int(transA), M, N, dt.type(alpha), dt.type(beta),
       np.ascontiguousarray(A.ravel()), np.ascontiguousarray(x.ravel()),
       _ptr(None if yold is None else np.ascontiguousarray(yold.ravel())), y)
    return y


def ref_saxpy(alpha, x, y):
    dt = x.dtype
    z = np.empty_like(x)
    getattr(lib(), "ora_ref_saxpy_" + _sfx(dt))(x.size, dt.type(alpha), x, y, z)
    return z


def absdiff(a, b):
    return float(getattr(lib(), "ora_absdiff_" + _sfx(a.dtype))(a.ravel(), b.ravel(), a.size))


def paranoid_pass(a, b):
    """OCLAppUtil.hpp:212-237: absdiff < 1/length."""
    return bool(getattr(lib(), "ora_paranoid_pass_" + _sfx(a.dtype))(a.ravel(), b.ravel(), a.size))


def check_array(p, value):
    """OCLAppUtil.hpp:55-61: every element equals value exactly."""
    return bool(getattr(lib(), "ora_check_array_" + _sfx(p.dtype))(p.ravel(), p.size, p.dtype.type(value)))


def gemm64(A, B, M, N, K, transA=False, transB=False, alpha=1.0, beta=0.0, C0=None, rows=None,
           with_mag=True):
    """FP64 product alpha*op(A)op(B) + beta*C0 (row-major storage, leading
    dimensions implied by the transpose flags).  `rows` selects output rows.
    Returns (C64, MAG) shaped (len(rows) or M, N)."""
    lda = M if transA else K
    ldb = K if transB else N
    nrows = M if rows is None else len(rows)
    rows_arr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    C64 = np.empty(nrows * N, dtype=np.float64)
    MAG = np.empty(nrows * N, dtype=np.float64) if with_mag else None
    c0 = None if C0 is None else np.ascontiguousarray(C0.ravel())
    if A.dtype == np.float32 and B.dtype == np.float32:
        if c0 is not None:
            c0 = c0.astype(np.float32, copy=False)
        lib().ora_gemm_f64_from_f32(int(transA), int(transB), M, N, K, alpha, A.ravel(), lda,
                                    B.ravel(), ldb, beta, _ptr(c0), N, _ptr(rows_arr), nrows,
                                    C64, _ptr(MAG))
    else:
        if c0 is not None:
            c0 = c0.astype(np.float64, copy=False)
        lib().ora_gemm_f64(int(transA), int(transB), M, N, K, alpha,
                           np.ascontiguousarray(A.ravel(), dtype=np.float64), lda,
                           np.ascontiguousarray(B.ravel(), dtype=np.float64), ldb, beta, _ptr(c0), N,
                           _ptr(rows_arr), nrows, C64, _ptr(MAG))
    C64 = C64.reshape(nrows, N)
    return C64, (None if MAG is None else MAG.reshape(nrows, N))


def gemv64(A, x, M, N, transA=False, alpha=1.0, beta=0.0, y0=None):
    lda = M if transA else N
    y64 = np.empty(M, dtype=np.float64)
    MAG = np.empty(M, dtype=np.float64)
    lib().ora_gemv_f64_from_f32(int(transA), M, N, alpha, np.ascontiguousarray(A.ravel(), dtype=np.float32),
                                lda, np.ascontiguousarray(x, dtype=np.float32), beta,
                                _ptr(None if y0 is None else np.ascontiguousarray(y0, dtype=np.float32)),
                                y64, _ptr(MAG))
    return y64, MAG


def error_metrics(got, ref, mag=None):
    """(componentwise max |got-ref|/mag, normwise ||got-ref||/||ref||, max abs)."""
    got = np.ascontiguousarray(got.ravel())
    ref = np.ascontiguousarray(ref.ravel(), dtype=np.float64)
    out = np.zeros(3, dtype=np.float64)
    m = None if mag is None else np.ascontiguousarray(mag.ravel(), dtype=np.float64)
    getattr(lib(), "ora_error_metrics_" + _sfx(got.dtype))(got, ref, _ptr(m), got.size, out)
    return float(out[0]), float(out[1]), float(out[2])


def port_matmul_rows(A, B, C, M, N, K, row0, row1, gemm=False, alpha=1.0, beta=0.0,
                     transA=False, transB=False):
    """CPU port of the reference buffer-kernel arithmetic on rows [row0,row1)."""
    lib().ora_port_matmul_f32_rows(int(gemm), int(transA), int(transB), M, N, K,
                                   np.float32(alpha), np.float32(beta), A.ravel(), B.ravel(), C.ravel(),
                                   row0, row1)


def port_matvec_rows(A, x, y, N, row0, row1, gemv=False, alpha=1.0, beta=0.0):
    lib().ora_port_matvec_f32_rows(int(gemv), N, np.float32(alpha), np.float32(beta),
                                   A.ravel(), x, y, row0, row1)


def port_saxpy(alpha, x, y, z, i0=0, i1=None):
    """CPU port of the reference SAXPY arithmetic into a preallocated z."""
    lib().ora_port_saxpy_f32(i0, x.size if i1 is None else i1, np.float32(alpha), x, y, z)


def num_threads():
    return lib().ora_num_threads()
