// extern "C" layer of the grouped GEMM (persistent launch, HIP graph) and
// its per-size-bucket tile tuning.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <sstream>
#include <vector>

#include "../../include/gatlas.h"
#include "capi_internal.hpp"
#include "grouped_plan.hpp"
#include "host/bench.hpp"

using namespace gat;

struct gat_grouped {
    GroupedPlan plan;
};

static std::vector<GroupedDesc> to_descs(const gat_gemm_desc* d, size_t n) {
    std::vector<GroupedDesc> v(n);
    for (size_t i = 0; i < n; ++i)
        v[i] = GroupedDesc{d[i].transA ? 1 : 0, d[i].transB ? 1 : 0, (int64_t)d[i].M, (int64_t)d[i].N,
                           (int64_t)d[i].K, d[i].alpha, d[i].A, (int64_t)d[i].lda, d[i].B, (int64_t)d[i].ldb,
                           d[i].beta, (float*)d[i].C, (int64_t)d[i].ldc};
    return v;
}

extern "C" {

long gat_grouped_plan(int dtype, const gat_gemm_desc* descs, size_t count, const int* shapes, int* out,
                      size_t cap) {
    if ((count && !descs) || (cap && !out)) {
        set_error("gat_grouped_plan: null descriptors or output");
        return GAT_ERR_INVALID;
    }
    const std::vector<GroupedDesc> v = to_descs(descs, count);
    std::vector<GroupedProblemDev> probs;
    std::vector<int4> work;
    GroupedQueues q{};
    std::vector<int> used;
    double flops = 0;
    if (GroupedPlan::build_tables(nullptr, dtype, v.data(), v.size(), probs, work, q, used, flops, shapes) != ST_OK)
        return GAT_ERR_INVALID;
    for (size_t i = 0; i < work.size() && i < cap; ++i) {
        out[4 * i] = work[i].x;
        out[4 * i + 1] = work[i].y;
        out[4 * i + 2] = work[i].z;
        out[4 * i + 3] = work[i].w;
    }
    return (long)work.size();
}

gat_grouped* gat_grouped_create(gat_ctx* ctx, int dtype, const gat_gemm_desc* descs, size_t count, int use_graph) {
    if (!ctx || (count && !descs)) {
        set_error("gat_grouped_create: null context or descriptors");
        return nullptr;
    }
    std::unique_ptr<gat_grouped> g(new gat_grouped());
    const std::vector<GroupedDesc> v = to_descs(descs, count);
    if (g->plan.create(ctx->impl, dtype, v.data(), v.size(), use_graph != 0) != ST_OK) return nullptr;
    return g.release();
}

int gat_grouped_run(gat_grouped* g) {
    if (!g) {
        set_error("null grouped plan");
        return GAT_ERR_INVALID;
    }
    return g->plan.run();
}

int gat_grouped_info(gat_grouped* g, size_t* tiles, double* flops, int* grid) {
    if (!g) {
        set_error("null grouped plan");
        return GAT_ERR_INVALID;
    }
    if (tiles) *tiles = g->plan.tiles();
    if (flops) *flops = g->plan.flops();
    if (grid) *grid = g->plan.grid();
    return GAT_OK;
}

int gat_grouped_shape(gat_grouped* g, size_t i) { return g ? g->plan.shape(i) : -1; }

void gat_grouped_destroy(gat_grouped* g) { delete g; }

int gat_gemm_grouped(gat_ctx* ctx, int dtype, const gat_gemm_desc* descs, size_t count) {
    gat_grouped* g = gat_grouped_create(ctx, dtype, descs, count, 0);
    if (!g) return GAT_ERR_INVALID;
    int rc = g->plan.run();
    if (rc == GAT_OK && hipStreamSynchronize(ctx->impl.stream) != hipSuccess) rc = GAT_ERR_HIP;
    delete g;
    return rc;
}

int gat_grouped_tune(gat_ctx* ctx, int dtype, const gat_gemm_desc* descs, size_t count, int reps,
                     const char* journal_path) {
    if (!ctx || (count && !descs)) {
        set_error("gat_grouped_tune: null context or descriptors");
        return GAT_ERR_INVALID;
    }
    if (dtype != GAT_F32 && dtype != GAT_BF16 && dtype != GAT_BF16X3) {
        set_error("gat_grouped_tune: dtype must be f32, bf16 or bf16x3");
        return GAT_ERR_INVALID;
    }
    Context& c = ctx->impl;
    c.make_current();
    if (reps < 1) reps = 1;
    std::unique_ptr<Journal> journal;
    if (journal_path) {
        journal.reset(new Journal(journal_path));
        journal->loadMemo();
    }
    const char* form = dtype == GAT_F32 ? "groupedf32" : dtype == GAT_BF16 ? "groupedbf16" : "groupedbf16x3";
    std::map<BucketKey, size_t> rep;  // one representative problem per size bucket
    for (size_t i = 0; i < count; ++i) {
        GemmArgs a;
        a.transA = descs[i].transA ? 1 : 0;
        a.transB = descs[i].transB ? 1 : 0;
        a.M = (int64_t)descs[i].M;
        a.N = (int64_t)descs[i].N;
        a.K = (int64_t)descs[i].K;
        if (a.M && a.N) rep.emplace(bucket_of(dtype, a), i);
    }
    // A shape is judged inside a full persistent launch, as the grouped
    // kernel runs it: the representative is replicated until its 128 x 128
    // tiles cover kTilesPerLaunch, each replica writing its own slice of a
    // scratch output (the caller's C is never written).
    const size_t kTilesPerLaunch = 4 * 2 * (size_t)std::max(c.num_cu, 1);
    const size_t kScratchBytes = (size_t)256 << 20;
    int tuned = 0;
    for (const auto& kv : rep) {
        const gat_gemm_desc& d = descs[kv.second];
        const size_t tiles0 = ((d.M + 127) / 128) * ((d.N + 127) / 128);
        const size_t cbytes = d.M * d.N * sizeof(float);
        size_t r = (kTilesPerLaunch + tiles0 - 1) / tiles0;
        r = std::max<size_t>(1, std::min(r, kScratchBytes / cbytes));
        void* scratch = nullptr;
        int rc = c.workspace(r * cbytes, &scratch);
        if (rc) return rc;
        std::vector<GroupedDesc> reps_d(r);
        for (size_t i = 0; i < r; ++i)
            reps_d[i] = GroupedDesc{d.transA ? 1 : 0, d.transB ? 1 : 0, (int64_t)d.M, (int64_t)d.N, (int64_t)d.K,
                                    1.0, d.A, (int64_t)d.lda, d.B, (int64_t)d.ldb, 0.0,
                                    (float*)scratch + i * d.M * d.N, (int64_t)d.N};
        double best = 0;
        int best_shape = -1;
        std::vector<std::pair<std::vector<std::pair<int, int>>, std::vector<std::pair<int, int>>>> seen;
        for (int s = 0; s < GP_NSHAPES; ++s) {
            // shapes that tile this problem alike (e.g. 128 and 64 rows when
            // M <= 64) are timed once
            std::vector<std::pair<int, int>> rs, cs;
            GroupedPlan::tile_spans((int64_t)d.M, gp_shape_bm(s), rs, gp_edge32(GroupedPlan::form_of(dtype)));
            GroupedPlan::tile_spans((int64_t)d.N, gp_shape_bn(s), cs, gp_edge32(GroupedPlan::form_of(dtype)));
            bool dup = false;
            for (const auto& t : seen) dup = dup || (t.first == rs && t.second == cs);
            if (dup) continue;
            seen.emplace_back(rs, cs);
            std::ostringstream key;
            key << form << "_" << (d.transA ? 1 : 0) << "_" << (d.transB ? 1 : 0) << "_" << d.M << "_" << d.N << "_"
                << d.K << "_" << s << "_";
            double us = journal ? journal->memoTime(key.str(), 0) : -1;  // a journaled trial is not re-run
            if (us < 0) {
                if (journal) journal->takeMemo(key.str(), Journal::RUN_IN_PROGRESS);
                const std::vector<int> shapes(r, s);
                GroupedPlan plan;
                rc = plan.create(c, dtype, reps_d.data(), r, false, shapes.data());
                if (rc) return rc;
                rc = plan.run();  // warm-up
                if (rc) return rc;
                GAT_HIP_TRY(hipEventRecord(c.ev_start, c.stream));
                for (int k = 0; k < reps; ++k) {
                    rc = plan.run();
                    if (rc) return rc;
                }
                GAT_HIP_TRY(hipEventRecord(c.ev_stop, c.stream));
                GAT_HIP_TRY(hipEventSynchronize(c.ev_stop));
                float ms = 0.f;
                GAT_HIP_TRY(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
                us = std::max(1.0, std::round(1000.0 * ms / reps));
                if (journal) {
                    journal->takeMemo(key.str(), (int)us);
                    journal->takeMemo(key.str(), Journal::RUN_OK);
                }
            }
            if (best_shape < 0 || us < best) {
                best = us;
                best_shape = s;
            }
        }
        c.set_grouped_shape(kv.first, best_shape);
        ++tuned;
    }
    return tuned;
}

}  // extern "C"
