// Device descriptor of one problem of a grouped GEMM launch.
#pragma once

#include "../kernels.hpp"

namespace gat {

enum : int {
    GP_TA = 1,   // op(A) = A^T (A stored K x M)
    GP_TB = 2,   // op(B) = B^T (B stored N x K)
    GP_VEC = 4,  // 16-byte aligned operands and leading dimensions
};

// Tile shapes, all of four wavefronts of 32x32 blocks.  Interior shapes
// (2 x 2 waves, each TM x TN blocks): 0 = 128 x 128, 1 = 64 x 64,
// 2 = 128 x 64, 3 = 64 x 128; a problem has one.  Edge-only shapes: 4 =
// 128 x 32 (4 x 1 waves of one block) and 5 = 32 x 128 (1 x 4 waves).
// The last rows (columns) of a 128-wide shape are cut into the fewest
// strips of 128, 64 and 32 that cover them to a multiple of 32 (at most 16
// rows past the matrix for M a multiple of 16, against 48 with 64-wide
// strips alone); a 32-wide strip stands next to a 128-wide side, and
// where the other side is narrower it is widened to 64.
// Merged shapes of the BF16 form: 6 = 256 x 128 and 7 = 128 x 256 (2 x 2
// waves of 4 x 2 and 2 x 4 blocks), two neighbouring whole 128 x 128 tiles
// of a problem with K a multiple of 16 and aligned rows taken as one: a
// quarter fewer LDS and L2 bytes per product, and half the per-tile cost.
constexpr int GP_NSHAPES = 4;      // interior shapes
constexpr int GP_NSHAPES_ALL = 8;  // with the edge-only and merged ones
constexpr int gp_shape_bm(int s) {
    return s == 6 ? 256 : s == 5 ? 32 : (s == 0 || s == 2 || s == 4 || s == 7) ? 128 : 64;
}
constexpr int gp_shape_bn(int s) {
    return s == 7 ? 256 : s == 4 ? 32 : (s == 0 || s == 3 || s == 5 || s == 6) ? 128 : 64;
}
// shape of a bm x bn tile (a 32-wide side pairs with a 128-wide one, a
// 256-wide side with a 128-wide one)
constexpr int gp_shape_of(int bm, int bn) {
    return bm == 256 ? 6 : bn == 256 ? 7 : bm == 32 ? 5 : bn == 32 ? 4
         : bm == 128 ? (bn == 128 ? 0 : 2) : (bn == 128 ? 3 : 1);
}

struct GroupedProblemDev {
    const void* A;
    const void* B;
    float* C;
    int M, N, K;
    int flags;
    int64_t lda, ldb, ldc;
    float alpha, beta;
};

// The tile list is split into one queue per XCD (work-group i of the
// persistent grid runs on XCD i % 8): queue x holds entries
// [off[x], off[x+1]) of the list.
constexpr int GP_XCDS = 8;
struct GroupedQueues {
    int off[GP_XCDS + 1];
};

// Precision forms of a grouped launch (every problem of one launch shares it)
enum : int {
    GP_FORM_F32 = 0,   // FP32 operands, FP32 matrix instructions
    GP_FORM_BF16 = 1,  // bf16 operands, FP32 accumulation
    GP_FORM_X3 = 2     // FP32 operands split into bf16 hi/lo while staged, 3 products
};

// Whether a form cuts 128-wide edges into 32-wide strips (shapes 4 and 5):
// the split-BF16x3 routines of those strips raise the whole kernel's
// register spills (18 -> 59) and cost more than the edge work they save
// (224 -> 217 TFLOP/s on the 512-problem workload), so that form keeps the
// 64-wide strips alone.
constexpr bool gp_edge32(int form) { return form != GP_FORM_X3; }
// Whether a form merges pairs of 128 x 128 tiles into shapes 6 and 7 (the
// BF16 form: its 128 x 128 tile is bound by LDS traffic; the FP32 one is
// already bound by the matrix unit, and the split-BF16x3 routines of a
// 256-wide tile would not fit the registers).
constexpr bool gp_merge256(int form) { return form == GP_FORM_BF16; }

size_t grouped_smem_bytes(int form);
// Resident work-groups per CU of a form's persistent kernel (LDS-limited).
int grouped_wgs_per_cu(int form);
// Sets the kernel's dynamic LDS limit (once); needed before graph capture.
int grouped_gemm_prepare(int form);
// `work` holds int4 entries {problem, m0, n0, tile shape} in per-XCD queues `q`;
// `counter` points to GP_XCDS + 1 ints (one cursor per queue, one count of
// finished groups) that are zero before the first launch and zero again
// after every launch completes.
int launch_grouped_gemm(int form, const GroupedProblemDev* probs, const int4* work, const GroupedQueues& q,
                        int* counter, int grid, hipStream_t st);

}  // namespace gat
