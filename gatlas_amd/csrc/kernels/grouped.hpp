// Device descriptor of one problem of a grouped GEMM launch.
#pragma once

#include "../kernels.hpp"

namespace gat {

enum : int {
    GP_TA = 1,   // op(A) = A^T (A stored K x M)
    GP_TB = 2,   // op(B) = B^T (B stored N x K)
    GP_VEC = 4,  // 16-byte aligned operands and leading dimensions
};

// Tile shapes, all of four wavefronts (2 x 2, each wave TM x TN 32x32
// blocks): 0 = 128 x 128, 1 = 64 x 64, 2 = 128 x 64, 3 = 64 x 128.  A
// problem has one interior shape; where the last tile row or column holds
// 64 or fewer rows (columns) of a 128-wide shape, that strip takes the
// 64-wide shape of the same other side, so an edge computes at most 63
// rows or columns past the matrix instead of 127.
constexpr int GP_NSHAPES = 4;
constexpr int gp_shape_bm(int s) { return (s == 0 || s == 2) ? 128 : 64; }
constexpr int gp_shape_bn(int s) { return (s == 0 || s == 3) ? 128 : 64; }
constexpr int gp_shape_of(int bm, int bn) { return bm == 128 ? (bn == 128 ? 0 : 2) : (bn == 128 ? 3 : 1); }

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
