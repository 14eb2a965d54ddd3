// extern "C" layer: context, memory, timing and the direct entry points.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <unordered_set>

#include "../../include/gatlas.h"
#include "capi_internal.hpp"
#include "context.hpp"
#include "kernels.hpp"

namespace gat {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

hipError_t allow_dynamic_lds(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::set<std::pair<const void*, int>> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    if (done.count({kernel, dev})) return hipSuccess;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) done.insert({kernel, dev});
    return e;
}

}  // namespace gat

using namespace gat;


#define CTX_CHECK(ctx)                          \
    do {                                        \
        if (!(ctx)) {                           \
            set_error("null context");          \
            return GAT_ERR_INVALID;             \
        }                                       \
    } while (0)

extern "C" {

gat_ctx* gat_create(int device) {
    gat_ctx* c = new gat_ctx();
    if (c->impl.init(device) != ST_OK) {
        delete c;
        return nullptr;
    }
    return c;
}

void gat_destroy(gat_ctx* ctx) { delete ctx; }

const char* gat_last_error(void) { return last_error(); }

size_t gat_max_workgroup(gat_ctx* ctx) { return ctx ? (size_t)ctx->impl.max_workgroup : 0; }

int gat_device_info(gat_ctx* ctx, char* buf, size_t len) {
    CTX_CHECK(ctx);
    const std::string s = ctx->impl.info();
    if (buf && len) {
        std::strncpy(buf, s.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return GAT_OK;
}

void* gat_stream(gat_ctx* ctx) { return ctx ? (void*)ctx->impl.stream : nullptr; }

int gat_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gat_malloc(gat_ctx* ctx, size_t bytes, void** dptr) {
    CTX_CHECK(ctx);
    return ctx->impl.alloc(bytes, dptr);
}

int gat_free(gat_ctx* ctx, void* dptr) {
    CTX_CHECK(ctx);
    return ctx->impl.release(dptr);
}

int gat_memcpy_h2d(gat_ctx* ctx, void* dst, const void* src, size_t bytes) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->impl.stream));
    GAT_HIP_TRY(hipStreamSynchronize(ctx->impl.stream));
    return GAT_OK;
}

int gat_memcpy_d2h(gat_ctx* ctx, void* dst, const void* src, size_t bytes) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->impl.stream));
    GAT_HIP_TRY(hipStreamSynchronize(ctx->impl.stream));
    return GAT_OK;
}

int gat_memcpy_d2d(gat_ctx* ctx, void* dst, const void* src, size_t bytes) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->impl.stream));
    return GAT_OK;
}

int gat_memset(gat_ctx* ctx, void* dst, int value, size_t bytes) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->impl.stream));
    return GAT_OK;
}

int gat_sync(gat_ctx* ctx) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipStreamSynchronize(ctx->impl.stream));
    return GAT_OK;
}

int gat_timer_start(gat_ctx* ctx) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipEventRecord(ctx->impl.ev_start, ctx->impl.stream));
    return GAT_OK;
}

int gat_timer_stop(gat_ctx* ctx, float* ms) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    GAT_HIP_TRY(hipEventRecord(ctx->impl.ev_stop, ctx->impl.stream));
    GAT_HIP_TRY(hipEventSynchronize(ctx->impl.ev_stop));
    float t = 0.f;
    GAT_HIP_TRY(hipEventElapsedTime(&t, ctx->impl.ev_start, ctx->impl.ev_stop));
    if (ms) *ms = t;
    return GAT_OK;
}

int gat_fill_uniform(gat_ctx* ctx, int dtype, void* dptr, size_t n, uint64_t seed, uint64_t stream, double lo,
                     double hi) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    return launch_fill_uniform(dtype, dptr, n, seed, stream, lo, hi, ctx->impl.stream);
}

int gat_fill_normal(gat_ctx* ctx, int dtype, void* dptr, size_t n, uint64_t seed, uint64_t stream) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    return launch_fill_normal(dtype, dptr, n, seed, stream, ctx->impl.stream);
}

int gat_scale_columns(gat_ctx* ctx, float* dptr, size_t rows, size_t cols, uint64_t seed, uint64_t stream) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    return launch_scale_columns(dptr, rows, cols, seed, stream, ctx->impl.stream);
}

static GemmArgs make_args(int transA, int transB, size_t M, size_t N, size_t K, double alpha, const void* A,
                          size_t lda, size_t sA, const void* B, size_t ldb, size_t sB, double beta, void* C,
                          size_t ldc, size_t sC, size_t batch) {
    GemmArgs a;
    a.transA = transA ? 1 : 0;
    a.transB = transB ? 1 : 0;
    a.M = (int64_t)M;
    a.N = (int64_t)N;
    a.K = (int64_t)K;
    a.alpha = alpha;
    a.beta = beta;
    a.A = A;
    a.lda = (int64_t)lda;
    a.strideA = (int64_t)sA;
    a.B = B;
    a.ldb = (int64_t)ldb;
    a.strideB = (int64_t)sB;
    a.C = C;
    a.ldc = (int64_t)ldc;
    a.strideC = (int64_t)sC;
    a.batch = (int)batch;
    return a;
}

static int check_gemm_shape(const GemmArgs& a) {
    const int64_t rowsA = a.transA ? a.K : a.M, colsA = a.transA ? a.M : a.K;
    const int64_t rowsB = a.transB ? a.N : a.K, colsB = a.transB ? a.K : a.N;
    (void)rowsA;
    (void)rowsB;
    if (a.lda < (colsA > 0 ? colsA : 1) || a.ldb < (colsB > 0 ? colsB : 1) || a.ldc < (a.N > 0 ? a.N : 1)) {
        set_error("gemm: leading dimension smaller than the row length");
        return GAT_ERR_INVALID;
    }
    if ((a.M && a.N && !a.C) || (a.K && a.M && !a.A) || (a.K && a.N && !a.B)) {
        set_error("gemm: null operand");
        return GAT_ERR_INVALID;
    }
    if (a.batch < 0) {
        set_error("gemm: negative batch");
        return GAT_ERR_INVALID;
    }
    return GAT_OK;
}

int gat_gemm_strided(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, double alpha,
                     const void* A, size_t lda, size_t strideA, const void* B, size_t ldb, size_t strideB,
                     double beta, void* C, size_t ldc, size_t strideC, size_t batch) {
    CTX_CHECK(ctx);
    if (batch > (size_t)INT32_MAX) {
        set_error("gemm: batch above 2^31 - 1");
        return GAT_ERR_INVALID;
    }
    GemmArgs a = make_args(transA, transB, M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                           strideC, batch);
    int rc = check_gemm_shape(a);
    if (rc) return rc;
    ctx->impl.make_current();
    return ctx->impl.gemm(dtype, a, nullptr);
}

int gat_gemm(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, double alpha,
             const void* A, size_t lda, const void* B, size_t ldb, double beta, void* C, size_t ldc) {
    return gat_gemm_strided(ctx, dtype, transA, transB, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1);
}

int gat_gemm_cfg(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, double alpha,
                 const void* A, size_t lda, const void* B, size_t ldb, double beta, void* C, size_t ldc,
                 const gat_tile_cfg* cfg) {
    CTX_CHECK(ctx);
    GemmArgs a = make_args(transA, transB, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1);
    int rc = check_gemm_shape(a);
    if (rc) return rc;
    ctx->impl.make_current();
    MfmaCfg c;
    if (cfg) c = from_tile_cfg(*cfg);
    return ctx->impl.gemm(dtype, a, cfg ? &c : nullptr);
}

int gat_tile_cfg_count(int dtype) { return Context::num_cfgs(dtype); }

int gat_tile_cfg_get(int dtype, int i, gat_tile_cfg* out) {
    if (!out || i < 0 || i >= Context::num_cfgs(dtype)) {
        set_error("tile cfg index out of range");
        return GAT_ERR_INVALID;
    }
    MfmaCfg c = Context::cfg_at(dtype, i);
    *out = to_tile_cfg(c);
    return GAT_OK;
}

int gat_gemm_compat(gat_ctx* ctx, int dtype, int transA, int transB, size_t M, size_t N, size_t K, int gemm,
                    double alpha, const void* A, const void* B, double beta, void* C, size_t gs, size_t bh,
                    size_t vl, size_t packed) {
    CTX_CHECK(ctx);
    if (dtype != GAT_F32 && dtype != GAT_F64) {
        set_error("gemm compat: dtype must be f32 or f64");
        return GAT_ERR_INVALID;
    }
    if (!gemm) {
        alpha = 1.0;
        beta = 0.0;
    }
    GemmArgs a = make_args(transA, transB, M, N, K, alpha, A, transA ? M : K, M * K, B, transB ? K : N, K * N,
                           beta, C, N, M * N, packed);
    ctx->impl.make_current();
    CompatCfg c{(int)gs, (int)bh, (int)vl};
    return launch_gemm_compat(dtype, c, a, ctx->impl.stream);
}

int gat_gemv(gat_ctx* ctx, int dtype, int transA, size_t M, size_t N, double alpha, const void* A, size_t lda,
             const void* x, double beta, void* y) {
    CTX_CHECK(ctx);
    if ((M && !y) || (M && N && (!A || !x))) {
        set_error("gemv: null operand");
        return GAT_ERR_INVALID;
    }
    if (lda < (transA ? M : N)) {
        set_error("gemv: leading dimension smaller than the row length");
        return GAT_ERR_INVALID;
    }
    ctx->impl.make_current();
    void* work = nullptr;
    const size_t wb = gemv_workspace_bytes(dtype, transA, M, N);
    if (wb) {
        int rc = ctx->impl.workspace(wb, &work);
        if (rc) return rc;
    }
    return launch_gemv_native(dtype, transA, M, N, alpha, A, lda, x, beta, y, work, ctx->impl.stream);
}

int gat_gemv_noalloc(gat_ctx* ctx, int dtype, int transA, size_t M, size_t N, double alpha, const void* A,
                     size_t lda, const void* x, double beta, void* y) {
    CTX_CHECK(ctx);
    if ((M && !y) || (M && N && (!A || !x))) {
        set_error("gemv_noalloc: null operand");
        return GAT_ERR_INVALID;
    }
    if (lda < (transA ? M : N)) {
        set_error("gemv_noalloc: leading dimension smaller than the row length");
        return GAT_ERR_INVALID;
    }
    ctx->impl.make_current();
    return launch_gemv_noalloc(ctx->impl.device, dtype, transA ? 1 : 0, M, N, alpha, A, lda, x, beta, y,
                               ctx->impl.stream);
}

int gat_gemv_compat(gat_ctx* ctx, int dtype, int transA, size_t M, size_t N, int gemv, double alpha,
                    const void* A, const void* x, double beta, void* y, size_t gs, size_t bh, size_t vl,
                    size_t packed) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    return launch_gemv_compat(dtype, (int)vl, transA, M, N, gs, bh, (int)packed, gemv, alpha, beta, A, x, y,
                              ctx->impl.stream);
}

int gat_axpy(gat_ctx* ctx, int dtype, size_t n, double alpha, const void* x, const void* y, void* z) {
    CTX_CHECK(ctx);
    if (n && (!x || !y || !z)) {
        set_error("axpy: null operand");
        return GAT_ERR_INVALID;
    }
    ctx->impl.make_current();
    return launch_saxpy_native(dtype, n, alpha, x, y, z, ctx->impl.stream);
}

int gat_axpy_compat(gat_ctx* ctx, int dtype, size_t M, size_t N, double alpha, const void* x, const void* y,
                    void* z, size_t gH, size_t gW, size_t bh, size_t bw, size_t vl, size_t packed) {
    CTX_CHECK(ctx);
    ctx->impl.make_current();
    return launch_saxpy_compat(dtype, (int)vl, M, N, gH, gW, bh, bw, (int)packed, alpha, x, y, z,
                               ctx->impl.stream);
}

int gat_convert_f32_bf16(gat_ctx* ctx, const float* src, void* dst, size_t n) {
    CTX_CHECK(ctx);
    if (n && (!src || !dst)) {
        set_error("convert: null operand");
        return GAT_ERR_INVALID;
    }
    ctx->impl.make_current();
    return launch_convert_f32_bf16(src, dst, n, ctx->impl.stream);
}

int gat_quantize_fp8(gat_ctx* ctx, const float* src, size_t rows, size_t cols, size_t ld, int trans, void* out,
                     float* scales) {
    CTX_CHECK(ctx);
    if (!src || !out || !scales || rows > INT32_MAX || cols > INT32_MAX) {
        set_error("quantize fp8: invalid operand");
        return GAT_ERR_INVALID;
    }
    if (rows && cols && ld < (trans ? rows : cols)) {
        set_error("quantize fp8: leading dimension shorter than the stored rows");
        return GAT_ERR_INVALID;
    }
    ctx->impl.make_current();
    // element (r, c) = trans ? src[c*ld + r] : src[r*ld + c]
    return launch_quantize_fp8(src, trans ? 1 : (int64_t)ld, trans ? (int64_t)ld : 1, (int)rows, (int)cols, out,
                               scales, ctx->impl.stream);
}

int gat_gemm_fp8(gat_ctx* ctx, size_t M, size_t N, size_t K, double alpha, const void* A8, const float* scaleA,
                 const void* B8, const float* scaleB, double beta, float* C, size_t ldc) {
    CTX_CHECK(ctx);
    // resident e4m3 operands always come with their block scales (without
    // them the FP32-operand path would read the byte arrays as floats)
    if (M && N && K && (!A8 || !B8 || !scaleA || !scaleB)) {
        set_error("gemm_fp8: null operand or block scales");
        return GAT_ERR_INVALID;
    }
    GemmArgs a = make_args(0, 1, M, N, K, alpha, A8, K, 0, B8, K, 0, beta, C, ldc, 0, 1);
    a.scaleA = scaleA;
    a.scaleB = scaleB;
    int rc = check_gemm_shape(a);
    if (rc) return rc;
    ctx->impl.make_current();
    return ctx->impl.gemm(GAT_FP8, a, nullptr);
}

}  // extern "C"
