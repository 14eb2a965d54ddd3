// extern "C" layer of the grouped GEMV and the serving loop.
#include <memory>
#include <vector>

#include "../../include/gatlas.h"
#include "capi_internal.hpp"
#include "serve.hpp"

using namespace gat;

struct gat_server {
    Server impl;
};

extern "C" {

int gat_gemv_grouped(gat_ctx* ctx, const gat_gemv_desc* d, size_t count) {
    if (!ctx || (count && !d)) {
        set_error("gat_gemv_grouped: null context or descriptors");
        return GAT_ERR_INVALID;
    }
    if (count == 0) return GAT_OK;
    if (count > (size_t)INT32_MAX) {
        set_error("gat_gemv_grouped: too many problems");
        return GAT_ERR_INVALID;
    }
    Context& c = ctx->impl;
    c.make_current();
    std::vector<GemvProblemDev> probs(count);
    int64_t blocks = 0;
    for (size_t i = 0; i < count; ++i) {
        if ((d[i].M && d[i].N && (!d[i].A || !d[i].x)) || (d[i].M && !d[i].y) || d[i].lda < d[i].N ||
            d[i].M > (size_t)INT32_MAX * 4 || d[i].N > (size_t)INT32_MAX) {
            set_error("gat_gemv_grouped: invalid problem descriptor");
            return GAT_ERR_INVALID;
        }
        GemvProblemDev& g = probs[i];
        g.A = d[i].A;
        g.x = d[i].x;
        g.y = d[i].y;
        g.M = (int64_t)d[i].M;
        g.N = (int64_t)d[i].N;
        g.lda = (int64_t)d[i].lda;
        g.alpha = (float)d[i].alpha;
        g.beta = (float)d[i].beta;
        g.vec = ((uintptr_t)g.A % 16 == 0) && ((uintptr_t)g.x % 16 == 0) && (g.lda % 4 == 0);
        g.blk0 = (int)blocks;
        blocks += gemv_grouped_blocks(g.M);
        if (blocks > INT32_MAX) {
            set_error("gat_gemv_grouped: too many work-groups");
            return GAT_ERR_INVALID;
        }
    }
    void* dp = nullptr;
    int rc = c.workspace(count * sizeof(GemvProblemDev), &dp);
    if (rc) return rc;
    GAT_HIP_TRY(hipMemcpyAsync(dp, probs.data(), count * sizeof(GemvProblemDev), hipMemcpyHostToDevice, c.stream));
    rc = launch_grouped_gemv((const GemvProblemDev*)dp, (int)count, blocks, c.stream);
    if (rc) return rc;
    GAT_HIP_TRY(hipStreamSynchronize(c.stream));
    return GAT_OK;
}

void gat_serve_cfg_default(gat_serve_cfg* cfg) {
    if (!cfg) return;
    cfg->rate = 2000.0;
    cfg->duration = 1.0;
    cfg->gemm_fraction = 0.7;
    cfg->hbm_budget = (size_t)64000000000ull;  /* 64 GB */
    cfg->mode = GAT_SERVE_BATCHED;
    cfg->streams = 4;
    cfg->max_batch = 256;
    cfg->max_inflight = 4096;
    cfg->gemm_min = 64;
    cfg->gemm_max = 4096;
    cfg->gemv_min = 1024;
    cfg->gemv_max = 32768;
    cfg->keep = 0;
    cfg->keep_max = 0;
    cfg->seed = 1;
    cfg->gemm_dtype = GAT_F32;
}

long gat_serve_trace(const gat_serve_cfg* cfg, gat_serve_arrival* out, size_t cap) {
    if (!cfg) {
        set_error("gat_serve_trace: null configuration");
        return GAT_ERR_INVALID;
    }
    if (serve_check_cfg(*cfg)) return GAT_ERR_INVALID;
    std::vector<ServeReq> reqs;
    size_t pool = 0;
    serve_draw_trace(*cfg, reqs, pool);
    for (size_t i = 0; i < reqs.size() && i < cap && out; ++i) {
        out[i].t = reqs[i].t_arr;
        out[i].kind = reqs[i].kind;
        out[i].transA = reqs[i].ta;
        out[i].transB = reqs[i].tb;
        out[i].M = (size_t)reqs[i].M;
        out[i].N = (size_t)reqs[i].N;
        out[i].K = (size_t)reqs[i].K;
        out[i].footprint = reqs[i].foot;
    }
    return (long)reqs.size();
}

gat_server* gat_server_create(gat_ctx* ctx, const gat_serve_cfg* cfg) {
    if (!ctx || !cfg) {
        set_error("gat_server_create: null context or configuration");
        return nullptr;
    }
    std::unique_ptr<gat_server> s(new gat_server());
    if (s->impl.create(ctx->impl, *cfg) != ST_OK) return nullptr;
    return s.release();
}

int gat_server_run(gat_server* s, gat_serve_stats* stats) {
    if (!s) {
        set_error("null server");
        return GAT_ERR_INVALID;
    }
    return s->impl.run(stats);
}

size_t gat_server_kept(gat_server* s) { return s ? s->impl.kept() : 0; }

int gat_server_kept_get(gat_server* s, size_t i, gat_serve_req* out) {
    if (!s) {
        set_error("null server");
        return GAT_ERR_INVALID;
    }
    return s->impl.kept_get(i, out);
}

int gat_server_check_alone(gat_server* s, size_t* mismatches) {
    if (!s) {
        set_error("null server");
        return GAT_ERR_INVALID;
    }
    return s->impl.check_alone(mismatches);
}

void gat_server_destroy(gat_server* s) { delete s; }

}  // extern "C"
