// extern "C" layer of the parameter interface, the journal, the search and
// the autotuner.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <sstream>

#include "../../include/gatlas.h"
#include "capi_internal.hpp"
#include "host/bench.hpp"
#include "host/device_app.hpp"
#include "host/kernels_impl.hpp"
#include "host/search.hpp"

using namespace gat;


struct gat_kernel {
    std::unique_ptr<KernelInterface> k;
    gat_ctx* ctx = nullptr;
    std::unique_ptr<DeviceApp> app;  // lazily bound to ctx for device work
    std::string name, printed;
    DeviceApp* device() {
        if (!ctx) return nullptr;
        if (!app) app.reset(new DeviceApp(&ctx->impl));
        return app.get();
    }
};

struct gat_journal {
    std::unique_ptr<Journal> j;
};

static int bad(const char* m) {
    set_error(m);
    return GAT_ERR_INVALID;
}

static bool params_of(KernelInterface* k, std::vector<size_t>& p) {
    if (auto* m = as_matmul(k)) return m->getParams(p);
    if (auto* v = as_matvec(k)) return v->getParams(p);
    if (auto* s = as_saxpy(k)) return s->getParams(p);
    return false;
}

static size_t nparams(KernelInterface* k) {
    if (as_matmul(k)) return 12;
    return 9;
}

extern "C" {

gat_kernel* gat_kernel_new(gat_ctx* ctx, int family, int dtype, int vl, int variant) {
    KernelInterface* k = make_kernel(family, dtype, vl, variant);
    if (!k) {
        set_error("gat_kernel_new: no such kernel family / type / vector length / variant");
        return nullptr;
    }
    gat_kernel* h = new gat_kernel();
    h->k.reset(k);
    h->ctx = ctx;
    return h;
}

void gat_kernel_free(gat_kernel* k) { delete k; }

const char* gat_kernel_name(gat_kernel* k) {
    if (!k) return "";
    k->name = k->k->kernelName();
    return k->name.c_str();
}

int gat_kernel_param_count(gat_kernel* k) { return k ? (int)nparams(k->k.get()) : 0; }

int gat_kernel_set_params(gat_kernel* k, const size_t* params, size_t n) {
    if (!k || !params || n != nparams(k->k.get())) return bad("gat_kernel_set_params: wrong parameter count");
    k->k->setParams(std::vector<size_t>(params, params + n));
    return GAT_OK;
}

int gat_kernel_valid(gat_kernel* k) {
    std::vector<size_t> p;
    return k && params_of(k->k.get(), p) ? 1 : 0;
}

int gat_kernel_get_params(gat_kernel* k, size_t* out, size_t n) {
    std::vector<size_t> p;
    if (!k || !params_of(k->k.get(), p)) return bad("gat_kernel_get_params: parameters are not valid");
    if (n < p.size()) return bad("gat_kernel_get_params: output too small");
    std::memcpy(out, p.data(), p.size() * sizeof(size_t));
    return (int)p.size();
}

int gat_kernel_set_paranoid(gat_kernel* k) {
    if (!k) return bad("null kernel");
    k->k->paranoidCheck();
    return GAT_OK;
}

int gat_kernel_set_packed(gat_kernel* k, size_t packed) {
    if (!k) return bad("null kernel");
    if (auto* m = as_matmul(k->k.get())) m->setPackedCalc(packed);
    else if (auto* v = as_matvec(k->k.get())) v->setPackedCalc(packed);
    else if (auto* s = as_saxpy(k->k.get())) s->setPackedCalc(packed);
    return GAT_OK;
}

size_t gat_kernel_total_variations(gat_kernel* k) {
    if (!k) return 0;
    if (auto* m = as_matmul(k->k.get())) return m->totalVariations();
    if (auto* v = as_matvec(k->k.get())) return v->totalVariations();
    if (auto* s = as_saxpy(k->k.get())) return s->totalVariations();
    return 0;
}

size_t gat_kernel_flops(gat_kernel* k) { return k ? k->k->numberFlops() : 0; }

int gat_kernel_extra_detail(gat_kernel* k, size_t* out, size_t n) {
    if (!k) return bad("null kernel");
    const std::vector<size_t> d = k->k->extraParamDetail();
    for (size_t i = 0; i < d.size() && i < n; ++i) out[i] = d[i];
    return (int)d.size();
}

int gat_kernel_work_items(gat_kernel* k, size_t* global, size_t* nglobal, size_t* local, size_t* nlocal,
                          size_t cap) {
    if (!k) return bad("null kernel");
    const std::vector<size_t> g = k->k->globalWorkItems(), l = k->k->localWorkItems();
    for (size_t i = 0; i < cap; ++i) {
        global[i] = i < g.size() ? g[i] : 0;
        local[i] = i < l.size() ? l[i] : 0;
    }
    if (nglobal) *nglobal = g.size();
    if (nlocal) *nlocal = l.size();
    return GAT_OK;
}

const char* gat_kernel_print(gat_kernel* k) {
    if (!k) return "";
    std::ostringstream os;
    os << *k->k;
    k->printed = os.str();
    return k->printed.c_str();
}

int gat_kernel_build(gat_kernel* k) {
    DeviceApp* app = k ? k->device() : nullptr;
    if (!app) return bad("gat_kernel_build: kernel has no device context");
    return k->k->build(*app) ? GAT_OK : GAT_ERR_UNSUPPORTED;
}

int gat_kernel_set_args(gat_kernel* k, int sync_input) {
    DeviceApp* app = k ? k->device() : nullptr;
    if (!app) return bad("gat_kernel_set_args: kernel has no device context");
    return k->k->setArgs(*app, sync_input != 0) ? GAT_OK : GAT_ERR_HIP;
}

int gat_kernel_run(gat_kernel* k, size_t ntrials, double* usec) {
    DeviceApp* app = k ? k->device() : nullptr;
    if (!app) return bad("gat_kernel_run: kernel has no device context");
    Bench b(*app, *k->k, false);
    std::vector<size_t> p;
    if (!params_of(k->k.get(), p)) return bad("gat_kernel_run: parameters are not valid");
    const size_t us = b.run(ntrials, p, false, false, false);
    if (usec) *usec = (double)us;
    return us ? GAT_OK : GAT_ERR_UNSUPPORTED;
}

int gat_kernel_sync_output(gat_kernel* k) {
    DeviceApp* app = k ? k->device() : nullptr;
    if (!app) return bad("no device context");
    return k->k->syncOutput(*app) ? GAT_OK : GAT_ERR_HIP;
}

int gat_kernel_check(gat_kernel* k, int print, double* err) {
    DeviceApp* app = k ? k->device() : nullptr;
    if (!app) return bad("no device context");
    const bool ok = k->k->checkOutput(*app, print != 0);
    if (err) *err = k->k->lastCheckError();
    return ok ? 1 : 0;
}

// ---------------------------------------------------------------- journal
gat_journal* gat_journal_open(const char* path) {
    if (!path) return nullptr;
    gat_journal* j = new gat_journal();
    j->j.reset(new Journal(path));
    return j;
}

void gat_journal_close(gat_journal* j) { delete j; }

int gat_journal_load(gat_journal* j) { return j && j->j->loadMemo() ? GAT_OK : GAT_ERR_INVALID; }

int gat_journal_purge(gat_journal* j, int delete_times) { return j ? j->j->purgeMemo(delete_times != 0) : -1; }

size_t gat_journal_good(gat_journal* j) { return j ? j->j->memoGood() : 0; }

int gat_journal_state(gat_journal* j, gat_kernel* k, const size_t* params, size_t n) {
    if (!j || !k) return 0;
    return j->j->memoRunState(*k->k, std::vector<size_t>(params, params + n));
}

int gat_journal_time(gat_journal* j, gat_kernel* k, const size_t* params, size_t n, size_t trial) {
    if (!j || !k) return -1;
    return j->j->memoTime(*k->k, std::vector<size_t>(params, params + n), trial);
}

int gat_journal_take(gat_journal* j, gat_kernel* k, const size_t* params, size_t n, int value) {
    if (!j || !k) return bad("null journal or kernel");
    return j->j->takeMemo(*k->k, std::vector<size_t>(params, params + n), value) ? GAT_OK : GAT_ERR_INVALID;
}

// --------------------------------------------------------------- ranking
int gat_rank_bench(size_t nth, const int* ok, const double* avg, size_t n) {
    std::vector<bool> o(n);
    for (size_t i = 0; i < n; ++i) o[i] = ok[i] != 0;
    return AppUtil::rankBench(nth, o, std::vector<double>(avg, avg + n));
}

int gat_mark_bench(size_t topN, int* ok, const size_t* time, size_t n) {
    std::vector<bool> o(n);
    for (size_t i = 0; i < n; ++i) o[i] = ok[i] != 0;
    AppUtil::markBench(topN, o, std::vector<size_t>(time, time + n));
    for (size_t i = 0; i < n; ++i) ok[i] = o[i] ? 1 : 0;
    return GAT_OK;
}

int gat_mark_bench_avg(size_t topN, int* ok, const double* avg, size_t n) {
    std::vector<bool> o(n);
    for (size_t i = 0; i < n; ++i) o[i] = ok[i] != 0;
    AppUtil::markBenchAvg(topN, o, std::vector<double>(avg, avg + n));
    for (size_t i = 0; i < n; ++i) ok[i] = o[i] ? 1 : 0;
    return GAT_OK;
}

// --------------------------------------------------------- enumeration
int gat_search_params(gat_kernel* k, size_t max_workgroup, size_t vl, size_t max_block_height,
                      size_t max_group_size, int generalized, size_t M, size_t N, size_t K, int transA, int transB,
                      long group_size, long block_height, long extra_param, size_t* out, size_t max_out) {
    if (!k) return bad("null kernel");
    std::vector<std::vector<size_t>> pargs;
    KernelInterface* ki = k->k.get();
    const int var = kernel_variant(ki);
    if (auto* m = as_matmul(ki)) {
        if (var == VAR_MFMA) {
            std::vector<size_t> fixed;
            pargs = getParamsMatmulMfma(*m, generalized, M, N, K, transA, transB, fixed, extra_param);
        } else {
            pargs = getParamsMatmul(*m, max_workgroup, vl, max_block_height, max_group_size, generalized, M, N, K,
                                    transA, transB, group_size, block_height, extra_param);
        }
    } else if (auto* v = as_matvec(ki)) {
        if (var == VAR_NATIVE)
            pargs = getParamsMatvecNative(*v, generalized, M, N, transA, group_size, block_height, extra_param);
        else
            pargs = getParamsMatvec(*v, max_workgroup, vl, max_block_height, max_group_size, generalized, M, N,
                                    transA, group_size, block_height, extra_param);
    } else if (auto* s = as_saxpy(ki)) {
        // SAXPY: M, N are the matrix shape; transB carries the group width and
        // K the block width when the group / blocking is fixed.
        if (var == VAR_NATIVE)
            pargs = getParamsSaxpyNative(*s, M, N, group_size, block_height, extra_param);
        else
            pargs = getParamsSaxpy(*s, max_workgroup, vl, max_block_height, max_group_size, M, N, group_size,
                                   group_size == FREE ? FREE : (long)transB, block_height,
                                   block_height == FREE ? FREE : (long)K, extra_param);
    }
    const size_t np = nparams(ki);
    size_t c = 0;
    for (const auto& p : pargs) {
        if ((c + 1) * np > max_out) break;
        std::memcpy(out + c * np, p.data(), np * sizeof(size_t));
        ++c;
    }
    return (int)pargs.size();
}

int gat_search_em(gat_kernel* k, const char* journal_path, size_t max_workgroup, size_t vl,
                  size_t max_block_height, size_t max_group_size, int generalized, size_t M, size_t N, size_t K,
                  int transA, int transB, size_t ntrials, int paranoid, size_t* best, size_t nbest) {
    if (!k || !journal_path) return bad("gat_search_em: null kernel or journal path");
    // Without a context the search replays the journal only; a candidate
    // that would need a device run fails (and is journaled as such).
    std::unique_ptr<DeviceApp> local;
    DeviceApp* app = k->device();
    if (!app) {
        local.reset(new DeviceApp((Context*)nullptr));
        app = local.get();
    }
    Journal journal(journal_path);
    Bench bench(*app, *k->k, journal, false);
    QuietReports quiet(true);  // a library call: the caller's stdout stays clean
    LoopOpts o;
    o.numberTrials = ntrials ? ntrials : 1;
    o.paranoid = paranoid != 0;
    SearchSpec s;
    s.maxWorkGroupSize = max_workgroup;
    s.vl = vl;
    s.maxBlockHeight = max_block_height;
    s.maxGroupSize = max_group_size;
    s.generalized = generalized != 0;
    s.tA = transA != 0;
    s.tB = transB != 0;
    s.M = M;
    s.N = N;
    s.K = K;
    const std::vector<size_t> b = emSearchKernel(*k->k, bench, journal, o, s);
    if (b.empty()) {
        set_error("gat_search_em: no good kernel");
        return GAT_ERR_UNSUPPORTED;
    }
    for (size_t i = 0; i < b.size() && i < nbest; ++i) best[i] = b[i];
    return (int)b.size();
}

// -------------------------------------------------------------- autotuner
int gat_tune_gemm(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                  const char* journal_path, int verbose, gat_tile_cfg* best_out, double* gflops) {
    if (!ctx || !journal_path) return bad("gat_tune_gemm: null context or journal path");
    std::unique_ptr<KernelInterface> k(make_kernel(FAM_MATMUL, dtype, 1, VAR_MFMA));
    if (!k) return bad("gat_tune_gemm: precision form has no matrix-core kernel");
    KernelBaseMatmul* m = as_matmul(k.get());
    DeviceApp app(&ctx->impl);
    QuietReports quiet(verbose == 0);
    Journal journal(journal_path);
    Bench bench(app, *k, journal, verbose != 0);
    // about 3 ms of work per timed trial at an assumed 100 TFLOP/s, 1..64 launches
    const double flops = 2.0 * (double)M * (double)N * (double)K;
    bench.setLaunchesPerTrial((size_t)std::max(1.0, std::min(64.0, std::ceil(3e-3 * 100e12 / std::max(flops, 1.0)))));
    LoopOpts o;
    // The instantiated tile table is small, so the first step times every
    // entry (tile shape and BK/BLK variant alike); the second re-ranks the
    // winner's variants from the journal and confirms the fixed point.
    auto enumerate = [&](int step, const std::vector<size_t>& b) {
        if (step == 0) return getParamsMatmulMfma(*m, false, M, N, K, transA, transB, std::vector<size_t>(), FREE);
        return getParamsMatmulMfma(*m, false, M, N, K, transA, transB, std::vector<size_t>{b[7], b[8], b[9], b[10]},
                                   FREE);
    };
    auto bump = [](std::vector<size_t>&) { return false; };
    const std::vector<size_t> best = emSearch(*k, bench, journal, o, {7, 8, 9, 10, 11}, enumerate, bump);
    if (best.empty()) return GAT_ERR_UNSUPPORTED;
    m->setParams(best);
    const std::vector<size_t> d = m->extraParamDetail();
    const MfmaCfg c{(int)best[7], (int)best[8], (int)best[9], (int)best[10], mfma_depth(dtype, d[0]), kMfmaBlk[d[1]],
                    kMfmaWK[d[2]]};
    GemmArgs a;
    a.transA = transA;
    a.transB = transB;
    a.M = (int64_t)M;
    a.N = (int64_t)N;
    a.K = (int64_t)K;
    ctx->impl.set_tuned(bucket_of(dtype, a), c);
    if (best_out) *best_out = to_tile_cfg(c);
    if (gflops) {
        journal.loadMemo();
        const int us = journal.memoTime(*k, best, 0);
        *gflops = us > 0 ? (double)m->numberFlops() / us / 1000.0 : 0.0;
    }
    return GAT_OK;
}

int gat_tuned_set(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                  const gat_tile_cfg* cfg) {
    if (!ctx || !cfg) return bad("null context or config");
    const MfmaCfg c = from_tile_cfg(*cfg);
    if (!Context::has_cfg(dtype, c)) return bad("gat_tuned_set: configuration not instantiated");
    GemmArgs a;
    a.transA = transA;
    a.transB = transB;
    a.M = (int64_t)M;
    a.N = (int64_t)N;
    a.K = (int64_t)K;
    ctx->impl.set_tuned(bucket_of(dtype, a), c);
    return GAT_OK;
}

int gat_tuned_get(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, gat_tile_cfg* out) {
    if (!ctx || !out) return bad("null context or output");
    GemmArgs a;
    a.transA = transA;
    a.transB = transB;
    a.M = (int64_t)M;
    a.N = (int64_t)N;
    a.K = (int64_t)K;
    MfmaCfg c;
    const bool tuned = ctx->impl.get_tuned(bucket_of(dtype, a), &c);
    if (!tuned) c = Context::default_cfg(dtype, a);
    *out = to_tile_cfg(c);
    return tuned ? 1 : 0;
}

int gat_tuned_clear(gat_ctx* ctx) {
    if (!ctx) return bad("null context");
    ctx->impl.clear_tuned();
    return GAT_OK;
}

}  // extern "C"
