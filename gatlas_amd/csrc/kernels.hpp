// Launcher declarations for every hand-written gfx950 kernel of the library.
#pragma once

#include "common.hpp"

namespace gat {

// One GEMM problem, row-major storage as the reference lays it out
// (KernelBaseMatmul.hpp:59-80 MatmulDataLayout: transposed == column-major):
//   op(A)[i][k] = transA ? A[k*lda + i] : A[i*lda + k]
//   op(B)[k][j] = transB ? B[j*ldb + k] : B[k*ldb + j]
//   C[i*ldc + j] = alpha * sum_k op(A)[i][k] op(B)[k][j] + beta * C[i*ldc + j]
// `batch` same-shape problems follow each other at the given strides
// (the reference's packedCalc, KernelMatmulBuffer.hpp:229-230).
// beta == 0 never reads C.
struct GemmArgs {
    int transA = 0, transB = 0;
    int64_t M = 0, N = 0, K = 0;
    const void* A = nullptr;
    int64_t lda = 0, strideA = 0;
    const void* B = nullptr;
    int64_t ldb = 0, strideB = 0;
    void* C = nullptr;
    int64_t ldc = 0, strideC = 0;
    double alpha = 1.0, beta = 0.0;
    int batch = 1;
    // FP8 only: per-128x128-block FP32 scales of op(A) stored M x K
    // (ceil(M/128) x ceil(K/128), row-major) and of op(B)^T stored N x K
    // (ceil(N/128) x ceil(K/128), row-major).
    const float* scaleA = nullptr;
    const float* scaleB = nullptr;
};

// MFMA tile configuration.  The work-group is waves_m x waves_n wavefronts;
// each wave owns tm x tn MFMA output blocks of `blk` x `blk` (32 or 16);
// the macro tile is (waves_m*tm*blk) x (waves_n*tn*blk) and the LDS K depth bk.
struct MfmaCfg {
    int waves_m, waves_n, tm, tn, bk, blk;
    int wk = 1;  // wave groups sharing each staged K tile (FP32 and FP64 tiles only)
};

// Reference-geometry ("compat") GEMM: work-group gs x gs work-items, each
// owning bh consecutive rows x vl consecutive columns, K staged through LDS
// gs*vl at a time (KernelMatmulBuffer.hpp:174-342).
struct CompatCfg {
    int gs, bh, vl;
};

// ---- fill / convert (fill.hip, convert.hip) ----
int launch_fill_uniform(int dtype, void* out, size_t n, uint64_t seed, uint64_t stream, double lo,
                        double hi, hipStream_t st);
int launch_fill_normal(int dtype, void* out, size_t n, uint64_t seed, uint64_t stream, hipStream_t st);
int launch_scale_columns(float* a, size_t rows, size_t cols, uint64_t seed, uint64_t stream,
                         hipStream_t st);

// ---- SAXPY (saxpy.hip) ----
// z = alpha*x + y, n elements, separate multiply and add roundings.
// Streaming launch shape: `wg` work-items per group, `unroll` 16-byte
// vectors per operand per work-item, non-temporal or cached accesses.
struct AxpyCfg {
    int wg, unroll, nontemporal;
};
bool axpy_cfg_ok(const AxpyCfg& c);
int launch_saxpy_native(int dtype, size_t n, double alpha, const void* x, const void* y, void* z,
                        hipStream_t st, const AxpyCfg* cfg = nullptr);
// Reference geometry (KernelSaxpyBuffer.hpp:143-225): 1-D when N == vl.
int launch_saxpy_compat(int dtype, int vl, size_t M, size_t N, size_t gH, size_t gW, size_t bh,
                        size_t bw, int packed, double alpha, const void* x, const void* y, void* z,
                        hipStream_t st);

// ---- GEMV (gemv.hip) ----
// y = alpha*op(A)*x + beta*y.  Native: wavefront-cooperative rows (row-major)
// or column-split partials with a deterministic second pass (transposed);
// `work` must hold gemv_workspace_bytes() bytes.
// Row-major launch shape: `wg` work-items per group, `rows` rows per
// wavefront, `unroll` 16-byte loads in flight per row per lane.
struct GemvCfg {
    int wg, rows, unroll;
};
bool gemv_cfg_ok(const GemvCfg& c);
size_t gemv_workspace_bytes(int dtype, int transA, size_t M, size_t N);
int launch_gemv_native(int dtype, int transA, size_t M, size_t N, double alpha, const void* A,
                       size_t lda, const void* x, double beta, void* y, void* work, hipStream_t st,
                       const GemvCfg* cfg = nullptr);
// Allocation-free form: the transposed partials go to a static per-device
// array of GEMV_SCRATCH_BYTES (its address taken by gemv_scratch_init, once
// per device, with that device current); shapes whose partials
// (gemv_partial_bytes) exceed it are rejected.  Same kernels, same results.
constexpr size_t GEMV_SCRATCH_BYTES = (size_t)16 << 20;
int gemv_scratch_init(int device);
size_t gemv_partial_bytes(int dtype, int transA, size_t M, size_t N);
int launch_gemv_noalloc(int device, int dtype, int transA, size_t M, size_t N, double alpha, const void* A,
                        size_t lda, const void* x, double beta, void* y, hipStream_t st);
// Reference geometry (KernelMatvecBuffer.hpp:180-325): each work-item owns bh
// rows, j-ascending unfused chain per row.
int launch_gemv_compat(int dtype, int vl, int transA, size_t M, size_t N, size_t gs, size_t bh,
                       int packed, int gemv, double alpha, double beta, const void* A, const void* x,
                       void* y, hipStream_t st);

// ---- GEMM ----
int launch_gemm_compat(int dtype, const CompatCfg& cfg, const GemmArgs& a, hipStream_t st);
int launch_gemm_mfma_f32(const MfmaCfg& cfg, const GemmArgs& a, hipStream_t st);
int launch_gemm_mfma_f64(const MfmaCfg& cfg, const GemmArgs& a, hipStream_t st);
bool mfma_f32_has(const MfmaCfg& cfg);
bool mfma_f64_has(const MfmaCfg& cfg);
int mfma_f32_num_cfgs();
MfmaCfg mfma_f32_cfg(int i);
int mfma_f64_num_cfgs();
MfmaCfg mfma_f64_cfg(int i);
bool compat_has(int dtype, const CompatCfg& cfg);

// BF16 forms (gemm_mfma_bf16.hip): dtype DT_BF16 (bf16 inputs) or
// DT_BF16X3 (FP32 inputs split into bf16 hi/lo, three products).
int launch_gemm_mfma_bf16(int dtype, const MfmaCfg& cfg, const GemmArgs& a, hipStream_t st);
// Tile tables per form (dtype DT_BF16 or DT_BF16X3).
int mfma_bf16_num_cfgs(int dtype);
MfmaCfg mfma_bf16_cfg(int dtype, int i);
bool mfma_bf16_has(int dtype, const MfmaCfg& cfg);
int launch_convert_f32_bf16(const float* x, void* y, size_t n, hipStream_t st);

// Split-BF16x3 through pre-split planes (gemm_x3_split.hip): op(A) and
// op(B)^T are split once into hi/lo bf16 planes in a workspace (rows padded
// to whole 256-row tiles, k to a multiple of 32), then a direct-to-LDS GEMM
// runs on the planes.  The tile table lists it as the split-BF16x3
// configuration (2, 4, 4, 2, 32, 32).
struct X3SplitDims {
    int64_t mpad = 0, npad = 0, kpad = 0;
    size_t bytes_a = 0, bytes_b = 0;  // workspace of each plane pair
};
bool x3_split_cfg(const MfmaCfg& c);
MfmaCfg x3_split_cfg_value();
X3SplitDims x3_split_dims(int64_t M, int64_t N, int64_t K);
// element (r, k) of the rows x K source is src[r*rs + k*cs]
int launch_x3_split(const float* src, int64_t rs, int64_t cs, int rows, int K, int64_t rows_pad, int64_t kpad,
                    void* dst, hipStream_t st);
int launch_x3_split_gemm(const X3SplitDims& d, const void* As, const void* Bs, float* C, int M, int N,
                         int64_t ldc, float alpha, float beta, hipStream_t st);

// FP8 e4m3 (gemm_fp8.hip): block quantisation of a rows x cols operand whose
// element (r, c) is src[r*rs + c*cs] into row-major bytes + block scales,
// and the scaled GEMM (A8 = op(A) as M x K, B8 = op(B)^T as N x K).
int launch_quantize_fp8(const float* src, int64_t rs, int64_t cs, int rows, int cols, void* out, float* scales,
                        hipStream_t st);
int launch_gemm_fp8(const MfmaCfg& cfg, const GemmArgs& a, hipStream_t st);
int mfma_fp8_num_cfgs();
MfmaCfg mfma_fp8_cfg(int i);
bool mfma_fp8_has(const MfmaCfg& cfg);

}  // namespace gat
