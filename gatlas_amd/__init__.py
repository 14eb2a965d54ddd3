"""MI355X (gfx950) GEMM / GEMV / SAXPY kernel families with the GATLAS
parameter interface.  The compute path is libgatlas_hip.so (hand-written HIP
kernels behind a C ABI, include/gatlas.h); this package only binds it."""
from . import _capi
from ._capi import GatError, load

__all__ = ["GatError", "load", "_capi"]
