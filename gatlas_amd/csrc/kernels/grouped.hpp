// Device descriptor of one problem of a grouped GEMM launch.
#pragma once

#include "../kernels.hpp"

namespace gat {

enum : int {
    GP_TA = 1,          // op(A) = A^T (A stored K x M)
    GP_TB = 2,          // op(B) = B^T (B stored N x K)
    GP_VEC = 4,         // 16-byte aligned operands and leading dimensions
    GP_SHAPE_SHIFT = 4  // bits 4..7: tile shape (0 = 128x128, 1 = 64x64)
};

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
// `work` holds int4 entries {problem, m0, n0, 0} in per-XCD queues `q`;
// `counter` points to GP_XCDS + 1 ints (one cursor per queue, one count of
// finished groups) that are zero before the first launch and zero again
// after every launch completes.
int launch_grouped_gemm(int form, const GroupedProblemDev* probs, const int4* work, const GroupedQueues& q,
                        int* counter, int grid, hipStream_t st);

}  // namespace gat
