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
 * dtype GAT_BF16X3: A, B, C FP32 (split inside the library: while the
 *   tiles are staged, or -- tile configuration (2,4,4,2,32,32), the
 *   default for large problems -- by a split pass into hi/lo planes in the
 *   context workspace; both give the same values).
 * dtype GAT_FP8: A, B, C FP32; A and B are block-quantised to e4m3 in the
 *   context workspace, then multiplied (see gat_gemm_fp8 for resident e4m3). */
int gat_gemm(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, double alpha,
             const void *A, size_t lda, const void *B, size_t ldb, double beta, void *C, size_t ldc);

/* Same, `batch` problems at element strides strideA/B/C (packedCalc,
 * KernelMatmulBuffer.hpp:229-230,339). */
int gat_gemm_strided(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                     double alpha, const void *A, size_t lda, size_t strideA, const void *B, size_t ldb,
                     size_t strideB, double beta, void *C, size_t ldc, size_t strideC, size_t batch);

/* MFMA tile configuration: waves_m x waves_n wavefronts per work-group, each
 * owning tm x tn blocks of blk x blk outputs; LDS K depth bk.  wk > 1
 * (FP32 only): wk groups of waves_m x waves_n wavefronts share each staged K
 * tile, each taking a run of its k steps, and their sums are added at the
 * end; 0 or 1 means no split. */
typedef struct gat_tile_cfg {
    int waves_m, waves_n, tm, tn, bk, blk;
    int wk;
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

/* FP32 -> BF16 (round to nearest even), n elements (BF16-input GEMM operands). */
int gat_convert_f32_bf16(gat_ctx *ctx, const float *src, void *dst, size_t n);
/* FP8 e4m3 block quantisation of a rows x cols operand whose element (r, c) is
 * trans ? src[c*ld + r] : src[r*ld + c]: writes rows x cols bytes row-major and
 * ceil(rows/128) x ceil(cols/128) FP32 block scales (powers of two >= amax/448).
 * ld below cols (trans: below rows) is rejected with GAT_ERR_INVALID. */
int gat_quantize_fp8(gat_ctx *ctx, const float *src, size_t rows, size_t cols, size_t ld, int trans, void *out,
                     float *scales);
/* C = alpha * (A8 .* sA) (B8 .* sB)^T + beta*C with A8 = op(A) as M x K e4m3 and
 * B8 = op(B)^T as N x K e4m3 (both from gat_quantize_fp8), FP32 accumulate.
 * gat_gemm with dtype GAT_FP8 takes FP32 operands and quantises them first. */
int gat_gemm_fp8(gat_ctx *ctx, size_t M, size_t N, size_t K, double alpha, const void *A8, const float *scaleA,
                 const void *B8, const float *scaleB, double beta, float *C, size_t ldc);

/* y = alpha*op(A)*x + beta*y; op(A) is M x N (A stored N x M when transA). */
int gat_gemv(gat_ctx *ctx, int dtype, int transA, size_t M, size_t N, double alpha, const void *A, size_t lda,
             const void *x, double beta, void *y);
/* Allocation-free gat_gemv.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * the same bits as gat_gemv; it never allocates, grows or frees memory.  The
 * transposed form keeps its slice partial sums in a static device array of
 * GAT_GEMV_NOALLOC_SCRATCH bytes per GPU (held until the process ends)
 * instead of the context workspace, and rejects with GAT_ERR_INVALID a shape
 * whose partials exceed it (about M <= 4M rows for FP32, 2M for FP64, fewer
 * for short-and-wide shapes); the row-major form needs no scratch.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (several contexts or streams at once) race. */
#define GAT_GEMV_NOALLOC_SCRATCH ((size_t)16 << 20)
int gat_gemv_noalloc(gat_ctx *ctx, int dtype, int transA, size_t M, size_t N, double alpha, const void *A,
                     size_t lda, const void *x, double beta, void *y);
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

/* ---------------------------------------------------------------------- */
/* Grouped GEMM (no reference counterpart; the reference's only batching   */
/* is packedCalc, same-shape problems in one kernel): many independent     */
/* problems of any shapes/transposes in ONE persistent launch, optionally  */
/* captured in a HIP graph.  dtype GAT_F32 (FP32 operands), GAT_BF16      */
/* (bf16 operands) or GAT_BF16X3 (FP32 operands split into bf16 hi/lo as   */
/* they are staged, three products: FP32-level error); C is FP32.          */
/* ---------------------------------------------------------------------- */
typedef struct gat_gemm_desc {
    int transA, transB;
    size_t M, N, K;
    double alpha;
    const void *A;
    size_t lda;
    const void *B;
    size_t ldb;
    double beta;
    void *C;
    size_t ldc;
} gat_gemm_desc;
typedef struct gat_grouped gat_grouped;
/* Builds the problem table and the cost-ordered tile list (device-resident;
 * tile shape per problem from the context's tuned size buckets). */
gat_grouped *gat_grouped_create(gat_ctx *ctx, int dtype, const gat_gemm_desc *descs, size_t count,
                                int use_graph);
/* One persistent launch (a graph replay when created with use_graph). */
int gat_grouped_run(gat_grouped *g);
/* Host-only dry run of gat_grouped_create's schedule, no device or context
 * needed (no counterpart in the reference: for callers that inspect or test
 * a grouped launch's tile list).  Interior shapes from `shapes` (count
 * entries, gat_grouped_shape numbering), or the untuned size rule when
 * NULL.  Writes up to cap entries {problem, m0, n0, tile shape} to out
 * (4 ints each) in launch order, the XCD queues one after the other, and
 * returns the number of tiles, or GAT_ERR_INVALID.  Tile shapes 4 and 5 are
 * 128x32 and 32x128 edge strips, 6 and 7 the merged bf16 256x128 and
 * 128x256 tiles; operand pointers are only tested for 16-byte alignment. */
long gat_grouped_plan(int dtype, const gat_gemm_desc *descs, size_t count, const int *shapes, int *out,
                      size_t cap);
int gat_grouped_info(gat_grouped *g, size_t *tiles, double *flops, int *grid);
/* Interior tile shape of problem i: 0 = 128x128, 1 = 64x64, 2 = 128x64,
 * 3 = 64x128 (M x N).  The last rows (columns) of a 128-wide shape run as
 * strips of 64 and, for dtype f32 and bf16, 32; for bf16, pairs of whole
 * 128x128 tiles of a problem with K a multiple of 16 and 16-byte aligned
 * rows run as one 256x128 or 128x256 tile.  None of this changes a
 * problem's bits (the same single-launch tile computes them). */
int gat_grouped_shape(gat_grouped *g, size_t i);
void gat_grouped_destroy(gat_grouped *g);
/* create + run + synchronise + destroy. */
int gat_gemm_grouped(gat_ctx *ctx, int dtype, const gat_gemm_desc *descs, size_t count);
/* Times the four tile shapes (gat_grouped_shape) for one problem per size
 * bucket, inside a full persistent launch of replicas of that problem
 * (reading its A and B, writing library scratch: the caller's C is not
 * touched), and installs the fastest in the context's grouped-shape table
 * (single-problem GEMMs keep their own tuned tiles); shapes that tile the
 * problem alike are timed once.  journal_path (may be NULL): every trial is
 * journaled as the reference's benchmark journal does (GatlasBenchmark.cpp:
 * 44-137 format, key "grouped<form>_<tA>_<tB>_<M>_<N>_<K>_<shape>_": -3, the
 * microseconds of one launch, -4), and a trial already journaled is not run
 * again.  Returns the number of buckets tuned. */
int gat_grouped_tune(gat_ctx *ctx, int dtype, const gat_gemm_desc *descs, size_t count, int reps,
                     const char *journal_path);

/* Grouped GEMV (FP32, A row-major M x N): y = alpha*A*x + beta*y for many
 * problems in one launch; each result is bit-identical to gat_gemv's. */
typedef struct gat_gemv_desc {
    size_t M, N;
    double alpha;
    const float *A;
    size_t lda;
    const float *x;
    double beta;
    float *y;
} gat_gemv_desc;
int gat_gemv_grouped(gat_ctx *ctx, const gat_gemv_desc *descs, size_t count);

/* ---------------------------------------------------------------------- */
/* Serving loop (no reference counterpart): a synthetic open-loop trace of */
/* GEMM and GEMV requests (Poisson arrivals) over device-resident          */
/* operands, admitted by HBM footprint and continuously batched onto       */
/* concurrent streams; reports latency percentiles and throughput.         */
/* ---------------------------------------------------------------------- */
#define GAT_SERVE_BATCHED 0 /* admitted requests grouped into persistent launches */
#define GAT_SERVE_LOOP 1    /* one launch per request (the per-problem loop)      */

typedef struct gat_serve_cfg {
    double rate;                  /* mean arrivals per second (Poisson)            */
    double duration;              /* seconds of arrivals                           */
    double gemm_fraction;         /* share of GEMM requests; the rest are GEMV     */
    size_t hbm_budget;            /* bytes of admitted footprint at once: every    */
                                  /* request counts its operands and its result;   */
                                  /* operands are views into one resident pool, so */
                                  /* their share is a working-set bound, results   */
                                  /* are allocated in an arena of this size        */
    int mode;                     /* GAT_SERVE_BATCHED or GAT_SERVE_LOOP           */
    int streams;                  /* device streams launches overlap on            */
    int max_batch;                /* problems per batched launch                   */
    int max_inflight;             /* admitted, unfinished requests at once         */
    int gemm_min, gemm_max;       /* GEMM M, N, K log-uniform, multiples of 16     */
    int gemv_min, gemv_max;       /* GEMV M, N log-uniform, multiples of 16        */
    int keep;                     /* keep every keep-th result (0: none) ...       */
    int keep_max;                 /* ... up to keep_max results                    */
    unsigned long long seed;
    int gemm_dtype;               /* GEMM requests: GAT_F32 or GAT_BF16X3 (FP32    */
                                  /* operands and results either way)             */
} gat_serve_cfg;

typedef struct gat_serve_stats {
    size_t requests, gemm_requests, gemv_requests, launches, kept;
    double wall_s;                /* first arrival to last completion              */
    double flops, bytes;          /* algorithmic flops, operand+result bytes       */
    double tflops, req_per_s;
    double p50_ms, p99_ms, max_ms, p50_gemm_ms, p99_gemm_ms, p50_gemv_ms, p99_gemv_ms;
    double mean_queue_ms;         /* arrival to admission                          */
    double mean_batch;            /* requests per launch                           */
    double peak_footprint;        /* bytes admitted at once, maximum               */
    double device_busy;           /* share of wall time with a launch in flight    */
    double blocked_ms;            /* time requests waited while admission was     */
                                  /* stopped by the budget, max_inflight or arena */
    double offered_tflops;        /* the trace's flops / its arrival duration      */
    double peak_output_bytes;     /* output-arena bytes held at once (the part of  */
                                  /* the footprint that is a real allocation)      */
} gat_serve_stats;

typedef struct gat_serve_req {
    int kind;                     /* 0 GEMM, 1 GEMV                                */
    int transA, transB, shape;    /* shape: GEMM tile shape (gat_grouped_shape)    */
    size_t M, N, K;               /* GEMV: K = 0                                   */
    const float *A;
    size_t lda;
    const float *B;               /* GEMV: x                                       */
    size_t ldb;
    float *C;                     /* GEMV: y                                       */
    size_t ldc;
    double latency_ms;
    size_t id;                    /* index of the request in the trace             */
} gat_serve_req;

typedef struct gat_serve_arrival {
    double t;                     /* arrival, seconds from the start of the trace  */
    int kind, transA, transB;     /* kind 0 GEMM, 1 GEMV                           */
    size_t M, N, K;               /* GEMV: K = 0                                   */
    size_t footprint;             /* admitted bytes: operands + result             */
} gat_serve_arrival;

typedef struct gat_server gat_server;
void gat_serve_cfg_default(gat_serve_cfg *cfg);
/* The trace a configuration replays (deterministic in cfg->seed, no device
 * needed): writes up to cap arrivals, returns the total count or <0. */
long gat_serve_trace(const gat_serve_cfg *cfg, gat_serve_arrival *out, size_t cap);
/* Allocates the resident operand pool and the output arena (hbm_budget
 * bytes), draws the trace; nothing is launched yet. */
gat_server *gat_server_create(gat_ctx *ctx, const gat_serve_cfg *cfg);
/* Replays the trace in real time; returns when every request finished. */
int gat_server_run(gat_server *s, gat_serve_stats *stats);
/* Kept results of the last run (device pointers valid until destroy). */
size_t gat_server_kept(gat_server *s);
int gat_server_kept_get(gat_server *s, size_t i, gat_serve_req *out);
/* Recomputes every kept request alone (single launch, same kernel and tile
 * shape) and counts results that differ in any bit. */
int gat_server_check_alone(gat_server *s, size_t *mismatches);
void gat_server_destroy(gat_server *s);

/* ---------------------------------------------------------------------- */
/* Parameter interface: the KernelInterface boundary                      */
/* (GatlasBenchmark.hpp:28-58) with the KernelBase{Matmul,Matvec,Saxpy}    */
/* parameter vectors.  Parameter functions need no device (ctx may be     */
/* NULL); build/set_args/run/check need a context.                        */
/* ---------------------------------------------------------------------- */

/* family GAT_MATMUL/MATVEC/SAXPY, dtype GAT_F32/F64 (MFMA also BF16, BF16X3,
 * FP8), vl 1/2/4, variant GAT_BUFFER/IMAGE/MFMA/NATIVE.  Replaces the
 * benchMatmul.cpp:460-485 kernel-type table.  NULL if no such kernel. */
gat_kernel *gat_kernel_new(gat_ctx *ctx, int family, int dtype, int vl, int variant);
void gat_kernel_free(gat_kernel *k);
/* KernelInterface::kernelName, e.g. "matmulbufferfloat4" (KernelMatmulBuffer.hpp:58-62). */
const char *gat_kernel_name(gat_kernel *k);
/* 12 for matmul, 9 for matvec and saxpy (KernelBase*::getParams). */
int gat_kernel_param_count(gat_kernel *k);
/* KernelBase*::setParams (KernelBaseMatmul.cpp:327-362). */
int gat_kernel_set_params(gat_kernel *k, const size_t *params, size_t n);
/* KernelBase*::validParams (KernelBaseMatmul.cpp:263-291): 1 valid, 0 not. */
int gat_kernel_valid(gat_kernel *k);
/* KernelBase*::getParams (KernelBaseMatmul.cpp:293-325): count written, or <0 if invalid. */
int gat_kernel_get_params(gat_kernel *k, size_t *out, size_t n);
/* KernelInterface::paranoidCheck. */
int gat_kernel_set_paranoid(gat_kernel *k);
/* setPackedCalc (KernelBaseMatmul.hpp:235-248). */
int gat_kernel_set_packed(gat_kernel *k, size_t packed);
/* MatmulExtraParameter::totalVariations (KernelBaseMatmul.cpp:200). */
size_t gat_kernel_total_variations(gat_kernel *k);
/* KernelBase*::numberFlops (KernelBaseMatmul.cpp:378-383). */
size_t gat_kernel_flops(gat_kernel *k);
/* KernelInterface::extraParamDetail: number of digits, digits into out. */
int gat_kernel_extra_detail(gat_kernel *k, size_t *out, size_t n);
/* globalWorkItems / localWorkItems (KernelBaseMatmul.cpp:364-376): up to cap
 * entries each, with their counts in *nglobal / *nlocal. */
int gat_kernel_work_items(gat_kernel *k, size_t *global, size_t *nglobal, size_t *local, size_t *nlocal,
                          size_t cap);
/* KernelInterface::print: describes the compiled HIP instantiation. */
const char *gat_kernel_print(gat_kernel *k);
/* Bench::rebuildProgram (GatlasBenchmark.cpp:160-175): resolves the instantiation. */
int gat_kernel_build(gat_kernel *k);
/* KernelInterface::setArgs (KernelMatmulBuffer.hpp:83-171). */
int gat_kernel_set_args(gat_kernel *k, int sync_input);
/* Bench::run(ntrials, ...) without a journal (GatlasBenchmark.cpp:178-289): usec, 0 on failure. */
int gat_kernel_run(gat_kernel *k, size_t ntrials, double *usec);
/* KernelInterface::syncOutput. */
int gat_kernel_sync_output(gat_kernel *k);
/* KernelInterface::checkOutput: 1 pass, 0 fail; err = absdiff or componentwise error. */
int gat_kernel_check(gat_kernel *k, int print, double *err);

/* ---------------------------------------------------------------------- */
/* Journal (GatlasBenchmark.hpp:62-95), byte-compatible text format:      */
/*   "<kernelName>_<p0>_..._<pn>_\t<value>\n"                              */
/* ---------------------------------------------------------------------- */
typedef struct gat_journal gat_journal;
gat_journal *gat_journal_open(const char *path);
void gat_journal_close(gat_journal *j);
/* Journal::loadMemo (GatlasBenchmark.cpp:56-73). */
int gat_journal_load(gat_journal *j);
/* Journal::purgeMemo (GatlasBenchmark.cpp:75-105): bad-key count, -1 on error. */
int gat_journal_purge(gat_journal *j, int delete_times);
/* Journal::memoGood. */
size_t gat_journal_good(gat_journal *j);
/* Journal::memoRunState (GatlasBenchmark.cpp:109-115): 0 missing, -1..-4 states. */
int gat_journal_state(gat_journal *j, gat_kernel *k, const size_t *params, size_t n);
/* Journal::memoTime (GatlasBenchmark.cpp:117-127): usec of a trial, -1 if absent. */
int gat_journal_time(gat_journal *j, gat_kernel *k, const size_t *params, size_t n, size_t trial);
/* Journal::takeMemo (GatlasBenchmark.cpp:129-137). */
int gat_journal_take(gat_journal *j, gat_kernel *k, const size_t *params, size_t n, int value);

/* AppUtil::rankBench / markBench (GatlasAppUtil.cpp:285-364); ok[] is 0/1. */
int gat_rank_bench(size_t nth, const int *ok, const double *avg, size_t n);
int gat_mark_bench(size_t topN, int *ok, const size_t *time, size_t n);
int gat_mark_bench_avg(size_t topN, int *ok, const double *avg, size_t n);

/* Candidate enumeration getParams (benchMatmul.cpp:217-304, benchMatvec.cpp:202-289,
 * benchSaxpy.cpp:158-275); -1 marks a free parameter.  Writes up to max_out/count
 * vectors, returns the total number of candidates.  For SAXPY, transB carries the
 * fixed group width and K the fixed block width. */
int gat_search_params(gat_kernel *k, size_t max_workgroup, size_t vl, size_t max_block_height,
                      size_t max_group_size, int generalized, size_t M, size_t N, size_t K, int transA, int transB,
                      long group_size, long block_height, long extra_param, size_t *out, size_t max_out);

/* The two-step EM search of benchMatmul.cpp:536-650 / benchMatvec.cpp:521-640 /
 * benchSaxpy.cpp:498-600 for kernel k, journaled to journal_path (trials already
 * in the journal are replayed, bad keys skipped).  Without a context (k created
 * with ctx NULL) only the journal is replayed.  Writes the best parameter vector;
 * returns its length, or <0. */
int gat_search_em(gat_kernel *k, const char *journal_path, size_t max_workgroup, size_t vl,
                  size_t max_block_height, size_t max_group_size, int generalized, size_t M, size_t N, size_t K,
                  int transA, int transB, size_t ntrials, int paranoid, size_t *best, size_t nbest);

/* Autotuner (no reference counterpart; replaces the exhaustive OpenCL search for
 * the matrix-core kernels): times every entry of the instantiated tile table for
 * one shape (each trial chains enough launches for about 3 ms of work), then
 * confirms the winner as the EM search's fixed point; journaled to journal_path,
 * result installed for the shape's size bucket. */
int gat_tune_gemm(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                  const char *journal_path, int verbose, gat_tile_cfg *best, double *gflops);
/* Install / query the tile choice of a size bucket (get returns 1 if tuned, 0 default). */
int gat_tuned_set(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                  const gat_tile_cfg *cfg);
int gat_tuned_get(gat_ctx *ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K,
                  gat_tile_cfg *out);
/* Forget every tuned choice (single-problem tiles and grouped shapes): all
 * shapes use the default tile rules again. */
int gat_tuned_clear(gat_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* GATLAS_H */
