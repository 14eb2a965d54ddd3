"""Python mirror of the reference's parameter interface and journal, bound to
the C ABI (the C++ classes live in libgatlas_hip.so).

Kernel(...) mirrors KernelInterface + KernelBase{Matmul,Matvec,Saxpy}
(GatlasBenchmark.hpp:28-58, KernelBaseMatmul.hpp:278-429); Journal mirrors
GatlasBenchmark.hpp:62-95; rank_bench / mark_bench mirror
GatlasAppUtil.cpp:285-364.
"""
import ctypes

from . import _capi
from ._capi import GatError, check, load

MATMUL, MATVEC, SAXPY = _capi.GAT_MATMUL, _capi.GAT_MATVEC, _capi.GAT_SAXPY
BUFFER, IMAGE, MFMA, NATIVE = _capi.GAT_BUFFER, _capi.GAT_IMAGE, _capi.GAT_MFMA, _capi.GAT_NATIVE


def _szarr(vals):
    return (ctypes.c_size_t * len(vals))(*[int(v) for v in vals])


class Kernel:
    """One kernel family instance (e.g. matmulbufferfloat4)."""

    def __init__(self, family, dtype=_capi.GAT_F32, vl=4, variant=BUFFER, ctx=None):
        self.lib = load()
        self.ctx = ctx
        self.h = self.lib.gat_kernel_new(ctx.h if ctx is not None else None, family, dtype, vl, variant)
        if not self.h:
            raise GatError(_capi.last_error())
        self.nparams = self.lib.gat_kernel_param_count(self.h)

    def __del__(self):
        try:
            if self.h:
                self.lib.gat_kernel_free(self.h)
                self.h = None
        except Exception:
            pass

    def kernelName(self):
        return self.lib.gat_kernel_name(self.h).decode()

    def setParams(self, params):
        check(self.lib.gat_kernel_set_params(self.h, _szarr(params), len(params)), "setParams")

    def validParams(self):
        return bool(self.lib.gat_kernel_valid(self.h))

    def getParams(self):
        out = (ctypes.c_size_t * 16)()
        n = self.lib.gat_kernel_get_params(self.h, out, 16)
        return list(out[:n]) if n > 0 else None

    def totalVariations(self):
        return self.lib.gat_kernel_total_variations(self.h)

    def numberFlops(self):
        return self.lib.gat_kernel_flops(self.h)

    def extraParamDetail(self):
        out = (ctypes.c_size_t * 8)()
        n = self.lib.gat_kernel_extra_detail(self.h, out, 8)
        return list(out[:n])

    def workItems(self):
        g, l = (ctypes.c_size_t * 4)(), (ctypes.c_size_t * 4)()
        ng, nl = ctypes.c_size_t(), ctypes.c_size_t()
        check(self.lib.gat_kernel_work_items(self.h, g, ctypes.byref(ng), l, ctypes.byref(nl), 4), "workItems")
        return list(g[:ng.value]), list(l[:nl.value])

    def globalWorkItems(self):
        return self.workItems()[0]

    def localWorkItems(self):
        return self.workItems()[1]

    def paranoidCheck(self):
        check(self.lib.gat_kernel_set_paranoid(self.h), "paranoidCheck")

    def setPackedCalc(self, n):
        check(self.lib.gat_kernel_set_packed(self.h, n), "setPackedCalc")

    def print(self):
        return self.lib.gat_kernel_print(self.h).decode()

    # ---- device side ----
    def build(self):
        return self.lib.gat_kernel_build(self.h) == 0

    def setArgs(self, sync_input=False):
        check(self.lib.gat_kernel_set_args(self.h, int(sync_input)), "setArgs")

    def run(self, ntrials=1):
        """Bench::run without a journal: microseconds, 0 on failure."""
        us = ctypes.c_double()
        self.lib.gat_kernel_run(self.h, ntrials, ctypes.byref(us))
        return us.value

    def syncOutput(self):
        check(self.lib.gat_kernel_sync_output(self.h), "syncOutput")

    def checkOutput(self, print_output=False):
        err = ctypes.c_double()
        ok = self.lib.gat_kernel_check(self.h, int(print_output), ctypes.byref(err))
        return bool(ok), err.value

    def search(self, M, N, K=0, transA=False, transB=False, generalized=False, group_size=-1, block_height=-1,
               extra_param=-1, max_workgroup=1024, vl=None, max_block_height=10, max_group_size=10, max_out=20000):
        """getParams candidate enumeration (benchMatmul.cpp:217-304 etc.)."""
        vl = vl if vl is not None else 4
        out = (ctypes.c_size_t * (max_out * self.nparams))()
        n = self.lib.gat_search_params(self.h, max_workgroup, vl, max_block_height, max_group_size, int(generalized),
                                       M, N, K, int(transA), int(transB), group_size, block_height, extra_param, out,
                                       max_out * self.nparams)
        n = min(n, max_out)
        return [list(out[i * self.nparams:(i + 1) * self.nparams]) for i in range(n)]


    def em_search(self, journal_path, M, N, K=0, transA=False, transB=False, generalized=False, vl=4,
                  max_workgroup=1024, max_block_height=10, max_group_size=10, ntrials=1, paranoid=False):
        """EM search (benchMatmul.cpp:536-650); replays the journal, runs the
        device only for candidates without a record.  Returns the best
        parameter vector or None."""
        out = (ctypes.c_size_t * 16)()
        n = self.lib.gat_search_em(self.h, journal_path.encode(), max_workgroup, vl, max_block_height,
                                   max_group_size, int(generalized), M, N, K, int(transA), int(transB), ntrials,
                                   int(paranoid), out, 16)
        return list(out[:n]) if n > 0 else None


class Journal:
    MISSING, BUILD_IN_PROGRESS, BUILD_OK, RUN_IN_PROGRESS, RUN_OK = 0, -1, -2, -3, -4

    def __init__(self, path):
        self.lib = load()
        self.h = self.lib.gat_journal_open(path.encode())

    def __del__(self):
        try:
            if self.h:
                self.lib.gat_journal_close(self.h)
                self.h = None
        except Exception:
            pass

    def loadMemo(self):
        return self.lib.gat_journal_load(self.h) == 0

    def purgeMemo(self, deleteTimes=False):
        return self.lib.gat_journal_purge(self.h, int(deleteTimes))

    def memoGood(self):
        return self.lib.gat_journal_good(self.h)

    def memoRunState(self, kernel, params):
        return self.lib.gat_journal_state(self.h, kernel.h, _szarr(params), len(params))

    def memoTime(self, kernel, params, trial):
        return self.lib.gat_journal_time(self.h, kernel.h, _szarr(params), len(params), trial)

    def takeMemo(self, kernel, params, value):
        return self.lib.gat_journal_take(self.h, kernel.h, _szarr(params), len(params), int(value)) == 0


def rank_bench(nth, ok, avg):
    n = len(ok)
    return load().gat_rank_bench(nth, (ctypes.c_int * n)(*ok), (ctypes.c_double * n)(*avg), n)


def mark_bench(topN, ok, times):
    n = len(ok)
    o = (ctypes.c_int * n)(*ok)
    load().gat_mark_bench(topN, o, _szarr(times), n)
    return list(o)


def mark_bench_avg(topN, ok, avg):
    n = len(ok)
    o = (ctypes.c_int * n)(*ok)
    load().gat_mark_bench_avg(topN, o, (ctypes.c_double * n)(*avg), n)
    return list(o)
