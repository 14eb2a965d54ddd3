#include "grouped_plan.hpp"

#include <algorithm>

namespace gat {

// Tile rows per band when a problem's tiles are queued (an XCD's 64
// resident groups then cover about an 8 x 8 square of tiles)
constexpr int GP_BAND = 8;

// Pairs of consecutive 128-wide spans wholly inside a side of `extent`
// become one 256-wide span.
static void merge_pairs(std::vector<std::pair<int, int>>& spans, int64_t extent) {
    std::vector<std::pair<int, int>> out;
    out.reserve(spans.size());
    for (size_t i = 0; i < spans.size(); ++i) {
        if (spans[i].second == 128 && i + 1 < spans.size() && spans[i + 1].second == 128 &&
            spans[i].first + 256 <= extent) {
            out.emplace_back(spans[i].first, 256);
            ++i;
        } else {
            out.push_back(spans[i]);
        }
    }
    spans.swap(out);
}

GroupedPlan::~GroupedPlan() {
    if (!ctx_) return;
    ctx_->make_current();
    hipStreamSynchronize(ctx_->stream);
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    if (d_probs_) hipFree(d_probs_);
    if (d_work_) hipFree(d_work_);
    if (d_counter_) hipFree(d_counter_);
}

// A single-problem configuration with the numerics of a grouped tile shape
// (32x32 blocks, no k-split): every shape accumulates each element's
// products in the same k order, so the 128 x 64 shapes map to the 128 x 128
// configuration, and a grouped problem's bits equal a single launch's.
MfmaCfg GroupedPlan::shape_cfg(int dtype, int shape) {
    const bool small = shape == 1;
    if (dtype == DT_BF16X3) return small ? MfmaCfg{2, 2, 1, 1, 32, 32} : MfmaCfg{2, 2, 2, 2, 16, 32};
    if (dtype == DT_BF16) return small ? MfmaCfg{2, 2, 1, 1, 64, 32} : MfmaCfg{2, 2, 2, 2, 32, 32};
    return small ? MfmaCfg{2, 2, 1, 1, 32, 32} : MfmaCfg{2, 2, 2, 2, 32, 32};
}

void GroupedPlan::tile_spans(int64_t extent, int T, std::vector<std::pair<int, int>>& spans, bool edge32) {
    spans.clear();
    for (int64_t o = 0; o < extent; o += T) {
        const int64_t rest = extent - o;
        if (T != 128 || rest > 96 || (!edge32 && rest > 64)) {
            spans.emplace_back((int)o, T);
        } else if (rest <= 32 && edge32) {
            spans.emplace_back((int)o, 32);
        } else {
            spans.emplace_back((int)o, 64);
            if (rest > 64) spans.emplace_back((int)o + 64, 32);  // edge32 only
        }
    }
}

int GroupedPlan::form_of(int dtype) {
    switch (dtype) {
        case DT_F32: return GP_FORM_F32;
        case DT_BF16: return GP_FORM_BF16;
        case DT_BF16X3: return GP_FORM_X3;
        default: return -1;
    }
}

int GroupedPlan::choose_shape(const Context* ctx, int dtype, const GroupedDesc& d) {
    GemmArgs a;
    a.transA = d.transA;
    a.transB = d.transB;
    a.M = d.M;
    a.N = d.N;
    a.K = d.K;
    int s = 0;
    if (ctx && ctx->get_grouped_shape(bucket_of(dtype, a), &s)) return s;
    // untuned: 128-wide sides where a side fills two of them, else 64
    return gp_shape_of(d.M >= 256 ? 128 : 64, d.N >= 256 ? 128 : 64);
}

int GroupedPlan::build_tables(const Context* ctx, int dtype, const GroupedDesc* d, size_t n,
                              std::vector<GroupedProblemDev>& probs, std::vector<int4>& work, GroupedQueues& q,
                              std::vector<int>& shapes, double& flops, const int* force_shapes) {
    if (form_of(dtype) < 0) {
        set_error("grouped gemm: dtype must be f32, bf16 or bf16x3");
        return ST_INVALID;
    }
    const bool bf16 = dtype == DT_BF16;  // bf16 operands (x3 takes FP32 operands)
    probs.resize(n);
    shapes.assign(n, 0);
    flops = 0;
    // A unit is a block of tiles of one problem: a problem of few tiles is
    // one unit, a larger one is cut into up to GP_XCDS blocks.
    struct Unit {
        double cost, kcost;
        int p, r0, r1, c0, c1;  // tile rows [r0, r1) x tile columns [c0, c1)
    };
    std::vector<Unit> units;
    // per problem: (offset, extent) of its tile rows and columns
    std::vector<std::vector<std::pair<int, int>>> rows, cols;
    rows.reserve(n);
    cols.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const GroupedDesc& g = d[i];
        if (g.M < 0 || g.N < 0 || g.K < 0 || g.M > INT32_MAX || g.N > INT32_MAX || g.K > INT32_MAX ||
            (g.M && g.N && !g.C) || (g.M && g.K && !g.A) || (g.N && g.K && !g.B) ||
            g.lda < (g.transA ? g.M : g.K) || g.ldb < (g.transB ? g.K : g.N) || g.ldc < g.N) {
            set_error("grouped gemm: invalid problem descriptor");
            return ST_INVALID;
        }
        const int s = force_shapes ? force_shapes[i] : choose_shape(ctx, dtype, g);
        shapes[i] = s;
        const int W = bf16 ? 8 : 4;
        const bool vec = ((uintptr_t)g.A % 16 == 0) && ((uintptr_t)g.B % 16 == 0) && g.lda % W == 0 &&
                         g.ldb % W == 0;
        GroupedProblemDev& p = probs[i];
        p.A = g.A;
        p.B = g.B;
        p.C = g.C;
        p.M = (int)g.M;
        p.N = (int)g.N;
        p.K = (int)g.K;
        p.flags = (g.transA ? GP_TA : 0) | (g.transB ? GP_TB : 0) | (vec ? GP_VEC : 0);
        p.lda = g.lda;
        p.ldb = g.ldb;
        p.ldc = g.ldc;
        p.alpha = (float)g.alpha;
        p.beta = (float)g.beta;
        if (s < 0 || s >= GP_NSHAPES) {
            set_error("grouped gemm: tile shape out of range");
            return ST_INVALID;
        }
        rows.emplace_back();
        cols.emplace_back();
        tile_spans(g.M, gp_shape_bm(s), rows.back(), gp_edge32(form_of(dtype)));
        tile_spans(g.N, gp_shape_bn(s), cols.back(), gp_edge32(form_of(dtype)));
        // BF16: pairs of whole 128 x 128 tiles along the longer side become
        // one 256-wide tile where its direct-to-LDS routine applies
        if (gp_merge256(form_of(dtype)) && s == 0 && vec && g.K >= 32 && g.K % 16 == 0)
            merge_pairs(g.M >= g.N ? rows.back() : cols.back(), g.M >= g.N ? g.M : g.N);
        const int tm = (int)rows.back().size(), tn = (int)cols.back().size();
        const double kc = (double)(g.K + 32);  // cost model: tile area times (K + 32)
        // A problem of many tiles is cut into up to GP_XCDS blocks of tiles
        // (one per 32 tiles), bm tile rows by bn tile columns: each block
        // reads its A panels once per block column and its B panels once
        // per block row, so of the cuts into that many blocks the one with
        // the fewest panel bytes is taken (a cut into row bands only would
        // have every band read all of B).
        int bm = 1, bn = 1;
        if (tm > 0 && tn > 0) {
            const int want = std::min(GP_XCDS, std::max(1, tm * tn / 32));
            double best = -1;
            for (int a = 1; a <= want; ++a)
                for (int b = 1; a * b <= want; ++b) {
                    if (a > tm || b > tn || a * b != want) continue;
                    const double bytes = (double)g.M * b + (double)g.N * a;
                    if (best < 0 || bytes < best) {
                        best = bytes;
                        bm = a;
                        bn = b;
                    }
                }
        }
        for (int bi = 0; bi < bm && tm > 0 && tn > 0; ++bi)
            for (int bj = 0; bj < bn; ++bj) {
                const int r0 = bi * tm / bm, r1 = (bi + 1) * tm / bm;
                const int c0 = bj * tn / bn, c1 = (bj + 1) * tn / bn;
                if (r1 <= r0 || c1 <= c0) continue;
                double area = 0;
                for (int r = r0; r < r1; ++r)
                    for (int c = c0; c < c1; ++c) area += (double)rows.back()[r].second * cols.back()[c].second;
                units.push_back(Unit{area * kc, kc, (int)i, r0, r1, c0, c1});
            }
        flops += 2.0 * (double)g.M * (double)g.N * (double)g.K;
    }
    // Units onto XCD queues, longest first onto the least-loaded queue, so
    // each problem's operand panels stay in one XCD's L2 where possible.
    std::stable_sort(units.begin(), units.end(), [](const Unit& a, const Unit& b) { return a.cost > b.cost; });
    struct Tile {
        double cost;
        int p, m0, n0, shape;
    };
    std::vector<Tile> lists[GP_XCDS];
    double load[GP_XCDS] = {0};
    for (const Unit& u : units) {
        int x = 0;
        for (int y = 1; y < GP_XCDS; ++y)
            if (load[y] < load[x]) x = y;
        load[x] += u.cost;
        auto emit = [&](int m0, int h, int n0, int wdt) {
            // a 32-wide strip needs a 128-wide other side; else it is
            // widened to 64 (it is the last strip: the extra rows or
            // columns lie past the matrix and are not written)
            const int bm = (h == 32 && wdt != 128) ? 64 : h;
            const int bn = (wdt == 32 && h != 128) ? 64 : wdt;
            lists[x].push_back(Tile{(double)bm * bn * u.kcost, u.p, m0, n0, gp_shape_of(bm, bn)});
        };
        // Inside the block, bands of GP_BAND tile rows, column-major in each
        // band: consecutive tiles share a B panel, and the groups of one
        // XCD (2 per CU x 32 CUs) working on a band at once read about
        // GP_BAND A and GP_BAND B panels, not the block's whole height of A.
        for (int rb = u.r0; rb < u.r1; rb += GP_BAND)
            for (int ci = u.c0; ci < u.c1; ++ci)
                for (int r = rb; r < std::min(rb + GP_BAND, u.r1); ++r) {
                    const auto& rw = rows[u.p][r];
                    const auto& c = cols[u.p][ci];
                    // a 256-wide span pairs with a whole 128-wide one only;
                    // next to an edge tile it is the two 128-wide tiles again
                    const bool whole_c = c.second == 128 && c.first + 128 <= d[u.p].N;
                    const bool whole_r = rw.second == 128 && rw.first + 128 <= d[u.p].M;
                    if (rw.second == 256 && !whole_c) {
                        emit(rw.first, 128, c.first, c.second);
                        emit(rw.first + 128, 128, c.first, c.second);
                    } else if (c.second == 256 && !whole_r) {
                        emit(rw.first, rw.second, c.first, 128);
                        emit(rw.first, rw.second, c.first + 128, 128);
                    } else {
                        emit(rw.first, rw.second, c.first, c.second);
                    }
                }
    }
    work.clear();
    for (int x = 0; x < GP_XCDS; ++x) {
        // within a queue the longest tiles start first; ties keep unit order
        std::stable_sort(lists[x].begin(), lists[x].end(),
                         [](const Tile& a, const Tile& b) { return a.cost > b.cost; });
        q.off[x] = (int)work.size();
        for (const Tile& t : lists[x]) work.push_back(make_int4(t.p, t.m0, t.n0, t.shape));
    }
    q.off[GP_XCDS] = (int)work.size();
    return ST_OK;
}

int GroupedPlan::create(Context& ctx, int dtype, const GroupedDesc* d, size_t n, bool use_graph,
                        const int* force_shapes) {
    std::vector<GroupedProblemDev> probs;
    std::vector<int4> work;
    int rc = build_tables(&ctx, dtype, d, n, probs, work, queues_, shapes_, flops_, force_shapes);
    if (rc) return rc;
    ctx_ = &ctx;
    form_ = form_of(dtype);
    ctx.make_current();
    nwork_ = work.size();
    const size_t pbytes = std::max<size_t>(1, probs.size()) * sizeof(GroupedProblemDev);
    const size_t wbytes = std::max<size_t>(1, work.size()) * sizeof(int4);
    GAT_HIP_TRY(hipMalloc(&d_probs_, pbytes));
    GAT_HIP_TRY(hipMalloc(&d_work_, wbytes));
    GAT_HIP_TRY(hipMalloc((void**)&d_counter_, (GP_XCDS + 1) * sizeof(int)));
    // the kernel leaves its counters at zero; they start there once
    GAT_HIP_TRY(hipMemsetAsync(d_counter_, 0, (GP_XCDS + 1) * sizeof(int), ctx.stream));
    if (!probs.empty())
        GAT_HIP_TRY(hipMemcpyAsync(d_probs_, probs.data(), probs.size() * sizeof(GroupedProblemDev),
                                   hipMemcpyHostToDevice, ctx.stream));
    if (!work.empty())
        GAT_HIP_TRY(hipMemcpyAsync(d_work_, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   ctx.stream));
    GAT_HIP_TRY(hipStreamSynchronize(ctx.stream));
    // the resident work-groups (LDS-limited per CU), never more than the tiles
    grid_ = std::max(1, std::min<int>(grouped_wgs_per_cu(form_) * (ctx.num_cu > 0 ? ctx.num_cu : 256),
                                      (int)std::max<size_t>(1, nwork_)));
    if (use_graph) {
        rc = grouped_gemm_prepare(form_);  // attribute calls stay outside the capture
        if (rc) return rc;
        GAT_HIP_TRY(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
        rc = launch_grouped_gemm(form_, (const GroupedProblemDev*)d_probs_, (const int4*)d_work_, queues_,
                                 d_counter_, grid_, ctx.stream);
        hipError_t e = hipStreamEndCapture(ctx.stream, &graph_);
        if (rc) return rc;
        GAT_HIP_TRY(e);
        GAT_HIP_TRY(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    }
    return ST_OK;
}

int GroupedPlan::run() {
    ctx_->make_current();
    if (exec_) {
        GAT_HIP_TRY(hipGraphLaunch(exec_, ctx_->stream));
        return ST_OK;
    }
    return launch_grouped_gemm(form_, (const GroupedProblemDev*)d_probs_, (const int4*)d_work_, queues_, d_counter_,
                               grid_, ctx_->stream);
}

}  // namespace gat
