// Host side of the persistent grouped GEMM: problem table, cost-ordered
// tile list, per-bucket tile choice, and the optional HIP graph holding
// the (counter reset, persistent kernel) pair.
#pragma once

#include <utility>
#include <vector>

#include "context.hpp"
#include "kernels/grouped.hpp"

namespace gat {

struct GroupedDesc {
    int transA, transB;
    int64_t M, N, K;
    double alpha;
    const void* A;
    int64_t lda;
    const void* B;
    int64_t ldb;
    double beta;
    float* C;
    int64_t ldc;
};

class GroupedPlan {
public:
    ~GroupedPlan();
    // dtype DT_F32 (FP32 operands), DT_BF16 (bf16 operands) or DT_BF16X3
    // (FP32 operands, split-BF16x3 products); FP32 output.
    // force_shapes (optional): the tile shape of every problem
    int create(Context& ctx, int dtype, const GroupedDesc* d, size_t n, bool use_graph,
               const int* force_shapes = nullptr);
    int run();  // one persistent launch (graph replay when captured)
    size_t tiles() const { return nwork_; }
    double flops() const { return flops_; }
    int grid() const { return grid_; }
    // interior tile shape of problem i (0 = 128x128, 1 = 64x64, 2 = 128x64,
    // 3 = 64x128); its edge strips may take a narrower shape
    int shape(size_t i) const { return shapes_[i]; }

    // shape for one problem: the context's tuned choice for its size bucket
    // when one exists (ctx may be null), else a size rule
    static int choose_shape(const Context* ctx, int dtype, const GroupedDesc& d);
    static MfmaCfg shape_cfg(int dtype, int shape);
    // (offset, extent) of the tiles of width T along a side of `extent`;
    // with T = 128 the last rows are cut into strips of 64 and (edge32) 32
    static void tile_spans(int64_t extent, int T, std::vector<std::pair<int, int>>& spans, bool edge32);
    // grouped-kernel precision form of a dtype (-1: not supported)
    static int form_of(int dtype);
    // Device descriptors and the cost-ordered (problem, m0, n0, shape) tile
    // list for n problems; also the interior shape per problem and the total
    // flops.  Host work only: ctx (may be null) supplies tuned shapes.
    static int build_tables(const Context* ctx, int dtype, const GroupedDesc* d, size_t n,
                            std::vector<GroupedProblemDev>& probs, std::vector<int4>& work, GroupedQueues& q,
                            std::vector<int>& shapes, double& flops, const int* force_shapes = nullptr);

private:
    Context* ctx_ = nullptr;
    int form_ = GP_FORM_F32;
    void* d_probs_ = nullptr;
    void* d_work_ = nullptr;
    int* d_counter_ = nullptr;
    size_t nwork_ = 0;
    GroupedQueues queues_{};
    int grid_ = 0;
    double flops_ = 0;
    std::vector<int> shapes_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace gat
