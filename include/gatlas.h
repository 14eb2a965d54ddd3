/*
 * gatlas.h -- C ABI of the MI355X (gfx950) GEMM / GEMV / SAXPY library.
 *
 * Plain C: opaque handles, pointers and sizes, no C++ types, no exceptions.
 * Every function returning int returns 0 on success and a negative status on
 * failure (GAT_ERR_*); gat_last_error() gives the message of the calling
 * thread's most recent failure.  Device pointers are HIP device addresses
 * (from gat_malloc or any HIP allocation on the context's device).  All work
 * of a context is issued in order on that context's single HIP stream, as the
 * reference serialises every trial on one queue (OCLBase.cpp:406-433,
 * GatlasBenchmark.cpp:221-236).
 *
 * Storage convention (KernelBaseMatmul.hpp:59-80, MatmulDataLayout): matrices
 * are row-major; a "transposed" operand is stored column-major, i.e. as the
 * row-major transpose:
 *   op(A)[i][k] = transA ? A[k*lda + i] : A[i*lda + k]
 *   op(B)[k][j] = transB ? B[j*ldb + k] : B[k*ldb + j]
 */
#ifndef GATLAS_H
#define GATLAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GAT_OK 0
#define GAT_ERR_INVALID (-1)
#define GAT_ERR_HIP (-2)
#define GAT_ERR_NOMEM (-3)
#define GAT_ERR_UNSUPPORTED (-4)

/* precision forms (the reference's float / double scalar types,
 * benchMatmul.cpp:141-160, plus the MFMA input forms) */
#define GAT_F32 0     /* FP32 in/out                                           */
#define GAT_F64 1     /* FP64 in/out, FP64 on the GPU                          */
#define GAT_BF16 2    /* BF16 inputs, FP32 accumulate and output               */
#define GAT_BF16X3 3  /* FP32 inputs split to BF16 hi+lo, 3 products, FP32 out */
#define GAT_FP8 4     /* FP8 e4m3 inputs + per-128x128-block FP32 scales       */

/* kernel families and variants of the parameter interface */
#define GAT_MATMUL 0
#define GAT_MATVEC 1
#define GAT_SAXPY 2
#define GAT_BUFFER 0  /* KernelMatmulBuffer / KernelMatvecBuffer / KernelSaxpyBuffer */
#define GAT_IMAGE 1   /* KernelMatmulImage / ... (same layout; interface compat)   */
#define GAT_MFMA 2    /* matrix-core kernels (no reference counterpart)            */
#define GAT_NATIVE 3  /* bandwidth kernels for matvec / saxpy (no counterpart)     */

typedef struct gat_ctx gat_ctx;
typedef struct gat_kernel gat_kernel;

/* ---------------------------------------------------------------------- */
/* Device context.  Replaces OCLBase + OCLApp (OCLBase.cpp:435-441,        */
/* OCLApp.hpp:31-330): one device, one in-order stream, events, buffers.   */
/* ---------------------------------------------------------------------- */

/* OCLApp::OCLApp(OCLBase&, size_t deviceIndex) -- OCLApp.hpp:84.  NULL on failure. */
gat_ctx *gat_create(int device);
/* OCLApp::~OCLApp -- releases every buffer the context owns. */
void gat_destroy(gat_ctx *ctx);
/* checkFail(...) messages (OCLUtil.tcc) -- message of the last failure on this thread. */
const char *gat_last_error(void);
/* OCLBase::maxWorkGroupSize (OCLBase.cpp:507). */
size_t gat_max_workgroup(gat_ctx *ctx);
/* oclInfo.cpp:27-69 analogue: one line of device properties. */
int gat_device_info(gat_ctx *ctx, char *buf, size_t len);
/* The context's HIP stream (hipStream_t) for interoperation. */
void *gat_stream(gat_ctx *ctx);
/* Number of GPUs visible to HIP. */
int gat_device_count(void);

/* device memory (OCLApp::createBuffer / enqueueWriteBuffer / enqueueReadBuffer,
 * OCLApp.tcc:127-174, OCLApp.cpp:417-448) */
int gat_malloc(gat_ctx *ctx, size_t bytes, void **dptr);
int gat_free(gat_ctx *ctx, void *dptr);
int gat_memcpy_h2d(gat_ctx *ctx, void *dst, const void *src, size_t bytes);
int gat_memcpy_d2h(gat_ctx *ctx, void *dst, const void *src, size_t bytes);
int gat_memcpy_d2d(gat_ctx *ctx, void *dst, const void *src, size_t bytes);
int gat_memset(gat_ctx *ctx, void *dst, int value, size_t bytes);
/* OCLApp::wait() (OCLApp.cpp:1000-1031): block until the stream drains. */
int gat_sync(gat_ctx *ctx);

/* Device timer over the context stream (hipEventRecord/ElapsedTime); the
 * reference times launch+wait with gettimeofday (GatlasBenchmark.cpp:209-272). */
int gat_timer_start(gat_ctx *ctx);
int gat_timer_stop(gat_ctx *ctx, float *ms);

/* Synthetic inputs generated on the device (no host transfer):
 * element i of stream s under seed: u = (splitmix64((seed*G ^ s<<48 ^ i) + G) >> 40) * 2^-24,
 * value = lo + (hi-lo)*u; G = 0x9E3779B97F4A7C15. */
int gat_fill_uniform(gat_ctx *ctx, int dtype, void *dptr, size_t n, uint64_t seed, uint64_t stream,
                     double lo, double hi);
int gat_fill_normal(gat_ctx *ctx, int dtype, void *dptr, size_t n, uint64_t seed, uint64_t stream);
/* Multiply column c of a rows x cols row-major FP32 matrix by 10^u_c, u_c ~ U(-3,3). */
int gat_scale_columns(gat_ctx *ctx, float *dptr, size_t rows, size_t cols, uint64_t seed, uint64_t stream);

/* ---------------------------------------------------------------------- */
/* Direct operand entry points (device pointers, asynchronous on the       */
/* context stream).  The reference's kernels compute these products in     */
/* KernelMatmulBuffer.hpp:174-342, KernelMatvecBuffer.hpp:180-325 and       */
/* KernelSaxpyBuffer.hpp:143-225; here the library picks the kernel.       */
/* ---------------------------------------------------------------------- */

/* C = alpha*op(A)*op(B) + beta*C; beta == 0 never reads C.
 * dtype GAT_F32 / GAT_F64: A, B, C of that type.
 * dtype GAT_BF16: A, B are bf16 arrays, C is FP32.
 * dtype GAT_BF16X3: A, B, C FP32 (split inside the library).
 * dtype GAT_FP8: A, B are e4m3 bytes of op(A) (M x K row-major) and op(B)
 *   stored as N x K row-major, scales via gat_gemm_fp8. */
int gat_gemm(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, double alpha,
             const void *A, size_t lda, const void *B, size_t ldb, double beta, void *C, size_t ldc);

/* Same, `batch` problems at element strides strideA/B/C (packedCalc,
 * KernelMatmulBuffer.hpp:229-230,339). */
int gat_gemm_strided(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                     double alpha, const void *A, size_t lda, size_t strideA, const void *B, size_t ldb,
                     size_t strideB, double beta, void *C, size_t ldc, size_t strideC, size_t batch);

/* MFMA tile configuration: waves_m x waves_n wavefronts per work-group, each
 * owning tm x tn blocks of blk x blk outputs; LDS K depth bk. */
typedef struct gat_tile_cfg {
    int waves_m, waves_n, tm, tn, bk, blk;
} gat_tile_cfg;

/* gat_gemm with an explicit MFMA tile configuration (used by the tuner). */
int gat_gemm_cfg(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, double alpha,
                 const void *A, size_t lda, const void *B, size_t ldb, double beta, void *C, size_t ldc,
                 const gat_tile_cfg *cfg);
/* Number / i-th of the instantiated tile configurations for a dtype. */
int gat_tile_cfg_count(int dtype);
int gat_tile_cfg_get(int dtype, int i, gat_tile_cfg *out);

/* Reference-geometry GEMM (compat engine): work-group gs x gs, blocking bh x vl,
 * bit-identical to the reference's host check (KernelMatmulBuffer.hpp:117-143).
 * gemm != 0 selects C = alpha*acc + beta*C, else C = acc.  Packed problems are
 * contiguous (strides M*K, K*N, M*N). */
int gat_gemm_compat(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, int gemm,
                    double alpha, const void *A, const void *B, double beta, void *C, size_t gs, size_t bh,
                    size_t vl, size_t packed);

/* y = alpha*op(A)*x + beta*y; op(A) is M x N (A stored N x M when transA). */
int gat_gemv(gat_ctx *ctx, int dtype, int transA, size_t M, size_t N, double alpha, const void *A, size_t lda,
             const void *x, double beta, void *y);
/* Reference-geometry GEMV (KernelMatvecBuffer.hpp:180-325), bit-identical to
 * the reference's host check (KernelMatvecBuffer.hpp:118-133). */
int gat_gemv_compat(gat_ctx *ctx, int dtype, int transA, size_t M, size_t N, int gemv, double alpha,
                    const void *A, const void *x, double beta, void *y, size_t gs, size_t bh, size_t vl,
                    size_t packed);

/* z = alpha*x + y over n elements (KernelSaxpyBuffer.hpp:143-225). */
int gat_axpy(gat_ctx *ctx, int dtype, size_t n, double alpha, const void *x, const void *y, void *z);
/* Reference geometry: M x N scalars, 1-D when N == vl (KernelBaseSaxpy.cpp:289-301). */
int gat_axpy_compat(gat_ctx *ctx, int dtype, size_t M, size_t N, double alpha, const void *x, const void *y,
                    void *z, size_t gH, size_t gW, size_t bh, size_t bw, size_t vl, size_t packed);

#ifdef __cplusplus
}
#endif

#endif /* GATLAS_H */
