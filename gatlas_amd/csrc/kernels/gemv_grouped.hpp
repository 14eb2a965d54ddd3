// Device descriptor and launcher of the grouped GEMV (FP32, row-major A).
#pragma once

#include "../kernels.hpp"

namespace gat {

// Per-wave shape of the grouped kernel (256 work-items per group): ROWS rows
// per wave, UNROLL loads in flight per row -- the rows/unroll of the
// single-problem default, so a problem's arithmetic is the same in both
// (the work-group size does not enter it).
enum : int { GEMV_GROUPED_ROWS = 2, GEMV_GROUPED_UNROLL = 1 };

struct GemvProblemDev {
    const float* A;
    const float* x;
    float* y;
    int64_t M, N, lda;
    float alpha, beta;
    int vec;   // 16-byte aligned A and x, lda % 4 == 0
    int blk0;  // first work-group of this problem (ascending over problems)
};

// work-groups one problem of M rows takes
int64_t gemv_grouped_blocks(int64_t M);
// `blocks` = sum of gemv_grouped_blocks over the n problems
int launch_grouped_gemv(const GemvProblemDev* probs, int n, int64_t blocks, hipStream_t st);

}  // namespace gat
