// Grouped GEMV: many independent y = alpha*A*x + beta*y problems (FP32,
// A row-major) in one launch.
//
// Problem p owns the work-groups [blk0_p, blk0_{p+1}); a work-group finds
// its problem by binary search over the ascending blk0 table, then runs
// the same wave routine with the same rows per wave and loads in flight as
// the single-problem kernel's default, so each problem's result is
// bit-identical to its single launch.
#include "gemv_grouped.hpp"
#include "gemv_rows.hpp"

namespace gat {

typedef float gv4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void grouped_gemv_kernel(const GemvProblemDev* __restrict__ probs, int n) {
    const int b = (int)blockIdx.x;
    int lo = 0, hi = n - 1;  // last problem whose blk0 <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (probs[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const GemvProblemDev& p = probs[lo];
    const int64_t wave = (int64_t)(b - p.blk0) * (256 / 64) + (threadIdx.x >> 6);
    const int64_t r0 = wave * GEMV_GROUPED_ROWS;
    if (r0 >= p.M) return;
    gemv_rows_wave<float, gv4, GEMV_GROUPED_ROWS, GEMV_GROUPED_UNROLL>(p.M, p.N, p.alpha, p.A, p.lda, p.x, p.beta,
                                                                      p.y, p.vec, r0, threadIdx.x & 63);
}

int64_t gemv_grouped_blocks(int64_t M) {
    const int64_t waves = (M + GEMV_GROUPED_ROWS - 1) / GEMV_GROUPED_ROWS;
    return (waves + 3) / 4;
}

int launch_grouped_gemv(const GemvProblemDev* probs, int n, int64_t blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return ST_OK;
    if (blocks > INT32_MAX) {
        set_error("grouped gemv: too many work-groups");
        return ST_INVALID;
    }
    hipLaunchKernelGGL(grouped_gemv_kernel, dim3((unsigned)blocks), dim3(256), 0, st, probs, n);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
