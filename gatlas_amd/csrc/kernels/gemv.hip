// GEMV:  y = alpha*op(A)*x + beta*y   (FP32 / FP64), HBM-bound.
//
// Native, A row-major (op(A) = A, M x N, row stride lda):
//   each wavefront owns ROWS consecutive rows (gemv_rows_wave, gemv_rows.hpp).
// Native, A transposed (A stored N x M row-major, op(A)[i][j] = A[j*lda + i]):
//   outputs are contiguous in memory, so a work-item owns 4 (FP32) or 2
//   (FP64) consecutive outputs and the j range is split into S slices across
//   work-groups.  Slice partial sums go to a workspace and a second kernel
//   adds the S partials in slice order (deterministic, no atomics) and
//   applies the epilogue.
// Compat (reference geometry, KernelMatvecBuffer.hpp:180-325): work-item g
//   owns rows g*bh .. g*bh+bh-1, each a j-ascending chain of separately
//   rounded multiply and add -- the arithmetic of the reference's host check.
#include <mutex>

#include "gemv_grouped.hpp"
#include "gemv_rows.hpp"

namespace gat {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename T, typename V, int ROWS, int UNROLL>
__global__ __launch_bounds__(512) void gemv_rows_kernel(int64_t M, int64_t N, T alpha, const T* __restrict__ A,
                                                         int64_t lda, const T* __restrict__ x, T beta,
                                                         T* __restrict__ y, int vec_ok) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t r0 = wave * ROWS;
    if (r0 >= M) return;
    gemv_rows_wave<T, V, ROWS, UNROLL>(M, N, alpha, A, lda, x, beta, y, vec_ok, r0, threadIdx.x & 63);
}

// Transposed: pass 1 partial sums.  A work-group of wg work-items covers
// wg*W consecutive outputs; blockIdx.y = slice of the j range.  Each
// work-item keeps JU rows of A in flight per step.
template <typename T, typename V, int JU>
__global__ __launch_bounds__(512) void gemv_cols_partial_kernel(int64_t M, int64_t N, const T* __restrict__ A,
                                                                 int64_t lda, const T* __restrict__ x,
                                                                 T* __restrict__ part, int64_t jslice,
                                                                 int vec_ok) {
    constexpr int W = sizeof(V) / sizeof(T);
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
    const int64_t j0 = (int64_t)blockIdx.y * jslice;
    int64_t j1 = j0 + jslice;
    if (j1 > N) j1 = N;
    if (i0 >= M) return;
    T* out = part + (int64_t)blockIdx.y * M;
    if (vec_ok && i0 + W <= M) {
        V acc = {};
        int64_t j = j0;
        for (; j + JU <= j1; j += JU) {
            V a[JU];
#pragma unroll
            for (int u = 0; u < JU; ++u) a[u] = gload_nt<V>(A + (j + u) * lda + i0);
#pragma unroll
            for (int u = 0; u < JU; ++u) acc += a[u] * gload<T>(x + j + u);
        }
        for (; j < j1; ++j) acc += gload<V>(A + j * lda + i0) * gload<T>(x + j);
        gstore(out + i0, acc);
    } else {
#pragma unroll
        for (int e = 0; e < W; ++e) {
            if (i0 + e >= M) break;
            T s = 0;
            for (int64_t j = j0; j < j1; ++j) s += A[j * lda + i0 + e] * x[j];
            out[i0 + e] = s;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_cols_reduce_kernel(int64_t M, int S, const T* __restrict__ part,
                                                                T alpha, T beta, T* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= M) return;
    T s = 0;
    for (int k = 0; k < S; ++k) s += part[(int64_t)k * M + i];
    T v = alpha * s;
    if (beta != (T)0) v += beta * y[i];
    y[i] = v;
}

// Slices of the j range for `target` work-groups in all, at least 64 rows
// of A each, at most 1024.
static size_t gemv_cols_max_slices(size_t N) {
    const size_t s = (N + 63) / 64;
    return s < 1 ? 1 : (s > 1024 ? 1024 : s);
}
static int gemv_cols_slices(size_t M, size_t N, int W, int wg, size_t target) {
    const size_t iblocks = (M + (size_t)wg * W - 1) / ((size_t)wg * W);
    size_t S = (target + iblocks - 1) / iblocks;
    const size_t maxS = gemv_cols_max_slices(N);
    if (S > maxS) S = maxS;
    if (S < 1) S = 1;
    return (int)S;
}

// The partial-sum workspace of the transposed form, for any launch shape.
size_t gemv_workspace_bytes(int dtype, int transA, size_t M, size_t N) {
    if (!transA) return 0;
    return gemv_cols_max_slices(N) * M * (dtype == DT_F64 ? 8 : 4);
}

template <typename T, typename V, int ROWS, int UNROLL>
static void launch_rows(size_t M, size_t N, int wg, T alpha, const T* A, size_t lda, const T* x, T beta, T* y,
                        int vec_ok, hipStream_t st) {
    const size_t waves = (M + ROWS - 1) / ROWS;
    const size_t wpb = (size_t)wg / 64;
    const size_t blocks = (waves + wpb - 1) / wpb;
    hipLaunchKernelGGL((gemv_rows_kernel<T, V, ROWS, UNROLL>), dim3((unsigned)blocks), dim3((unsigned)wg), 0, st,
                       (int64_t)M, (int64_t)N, alpha, A, (int64_t)lda, x, beta, y, vec_ok);
}

bool gemv_cfg_ok(const GemvCfg& c) {
    return (c.wg == 64 || c.wg == 128 || c.wg == 256 || c.wg == 512) &&
           (c.rows == 1 || c.rows == 2 || c.rows == 4 || c.rows == 8) &&
           (c.unroll == 1 || c.unroll == 2 || c.unroll == 4);
}

template <typename T, typename V>
static int launch_native_t(int transA, const GemvCfg& cfg, size_t M, size_t N, T alpha, const T* A, size_t lda,
                           const T* x, T beta, T* y, void* work, hipStream_t st) {
    constexpr int W = sizeof(V) / sizeof(T);
    if (!gemv_cfg_ok(cfg)) {
        set_error("gemv: unsupported (work-group, rows, unroll)");
        return ST_UNSUPPORTED;
    }
    if (!transA) {
        const int vec_ok = ((uintptr_t)A % 16 == 0) && ((uintptr_t)x % 16 == 0) && (lda % W == 0);
#define GAT_ROWS(R, U)                                                                              \
    if (cfg.rows == R && cfg.unroll == U)                                                           \
        launch_rows<T, V, R, U>(M, N, cfg.wg, alpha, A, lda, x, beta, y, vec_ok, st);
        GAT_ROWS(1, 1) GAT_ROWS(1, 2) GAT_ROWS(1, 4) GAT_ROWS(2, 1) GAT_ROWS(2, 2) GAT_ROWS(2, 4)
        GAT_ROWS(4, 1) GAT_ROWS(4, 2) GAT_ROWS(4, 4) GAT_ROWS(8, 1) GAT_ROWS(8, 2) GAT_ROWS(8, 4)
#undef GAT_ROWS
        GAT_HIP_TRY(hipGetLastError());
        return ST_OK;
    }
    if (!work) {
        set_error("gemv transposed: workspace required");
        return ST_INVALID;
    }
    // launch shape of the transposed form: cfg.wg work-items per group,
    // 2*cfg.rows rows of A in flight per work-item, 1024*cfg.unroll
    // work-groups in all (j range split into slices to reach them)
    const int S = gemv_cols_slices(M, N, W, cfg.wg, (size_t)1024 * cfg.unroll);
    const int64_t jslice = (int64_t)((N + S - 1) / S);
    // partial rows start at slice*M: vector stores need M % W == 0
    const int vec_ok = ((uintptr_t)A % 16 == 0) && (lda % W == 0) && (M % W == 0) &&
                       ((uintptr_t)work % 16 == 0);
    const size_t iblocks = (M + (size_t)cfg.wg * W - 1) / ((size_t)cfg.wg * W);
    const dim3 grid((unsigned)iblocks, (unsigned)S);
#define GAT_COLS(JU)                                                                                       \
    if (cfg.rows * 2 == JU)                                                                                \
        hipLaunchKernelGGL((gemv_cols_partial_kernel<T, V, JU>), grid, dim3((unsigned)cfg.wg), 0, st,       \
                           (int64_t)M, (int64_t)N, A, (int64_t)lda, x, (T*)work, jslice, vec_ok);
    GAT_COLS(2) GAT_COLS(4) GAT_COLS(8) GAT_COLS(16)
#undef GAT_COLS
    GAT_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((gemv_cols_reduce_kernel<T>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st,
                       (int64_t)M, S, (const T*)work, alpha, beta, y);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

// Default launch shapes: the fastest of the 16384^2 sweep.  Rows: 128
// work-items, 2 rows per wave, one load per row in flight (the grouped
// kernel uses the same rows/unroll so results match bit for bit).
// Transposed: 64 work-items (256 outputs) per group, 8 rows of A in flight
// per work-item, 1024 work-groups.
static constexpr GemvCfg kColsDefault{64, 4, 1};
static constexpr GemvCfg kRowsDefault{128, GEMV_GROUPED_ROWS, GEMV_GROUPED_UNROLL};

int launch_gemv_native(int dtype, int transA, size_t M, size_t N, double alpha, const void* A, size_t lda,
                       const void* x, double beta, void* y, void* work, hipStream_t st, const GemvCfg* cfg) {
    if (M == 0) return ST_OK;
    const GemvCfg c = cfg ? *cfg : transA ? kColsDefault : kRowsDefault;
    if (dtype == DT_F32)
        return launch_native_t<float, f4v>(transA, c, M, N, (float)alpha, (const float*)A, lda, (const float*)x,
                                           (float)beta, (float*)y, work, st);
    if (dtype == DT_F64)
        return launch_native_t<double, d2v>(transA, c, M, N, alpha, (const double*)A, lda, (const double*)x, beta,
                                            (double*)y, work, st);
    set_error("gemv: dtype must be f32 or f64");
    return ST_INVALID;
}

// ---------------- allocation-free entry (static scratch) ----------------
// The transposed form's slice partials, for callers that cannot allocate:
// one file-scope array per GPU, held until the process ends.  Host code
// never passes `scratch` itself (its host-side name is a stand-in, not a
// device address); each device's address comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address
// reaches a kernel.  Every partial the second pass reads is written by the
// first pass of the same call, so nothing needs clearing between calls;
// calls that overlap on one GPU (two streams) share the array and race.
static __device__ __align__(256) unsigned char scratch[GEMV_SCRATCH_BYTES];

static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices];

int gemv_scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) {
        set_error("gemv scratch: device index out of range");
        return ST_INVALID;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return ST_OK;
    void* p = nullptr;
    GAT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("gemv scratch: hipGetSymbolAddress returned a null address");
        return ST_HIP;
    }
    g_scratch_ptr[device] = p;
    return ST_OK;
}

size_t gemv_partial_bytes(int dtype, int transA, size_t M, size_t N) {
    if (!transA || M == 0) return 0;
    const int W = dtype == DT_F64 ? 2 : 4;
    const int S = gemv_cols_slices(M, N, W, kColsDefault.wg, (size_t)1024 * kColsDefault.unroll);
    return (size_t)S * M * (dtype == DT_F64 ? 8 : 4);
}

int launch_gemv_noalloc(int device, int dtype, int transA, size_t M, size_t N, double alpha, const void* A,
                        size_t lda, const void* x, double beta, void* y, hipStream_t st) {
    if (dtype != DT_F32 && dtype != DT_F64) {
        set_error("gemv: dtype must be f32 or f64");
        return ST_INVALID;
    }
    void* work = nullptr;
    const size_t need = gemv_partial_bytes(dtype, transA, M, N);
    if (need) {
        if (need > GEMV_SCRATCH_BYTES) {
            set_error("gemv_noalloc: the transposed form needs " + std::to_string(need) +
                      " bytes of partial sums, above the static scratch of " +
                      std::to_string((size_t)GEMV_SCRATCH_BYTES) + " bytes");
            return ST_INVALID;
        }
        std::lock_guard<std::mutex> g(g_scratch_mu);
        work = (device >= 0 && device < kMaxScratchDevices) ? g_scratch_ptr[device] : nullptr;
        if (!work) {
            set_error("gemv_noalloc: static scratch not set up on this device");
            return ST_INVALID;
        }
    }
    return launch_gemv_native(dtype, transA, M, N, alpha, A, lda, x, beta, y, work, st);
}

// ---------------- reference geometry ----------------

// acc + a*b and alpha*acc + beta*c, each product rounded before the add
// (contraction switched off: HIP device code defaults to fp-contract=fast).
template <typename T>
__device__ __forceinline__ T mad_rn(T acc, T a, T b) {
#pragma clang fp contract(off)
    const T p = a * b;
    return acc + p;
}
template <typename T>
__device__ __forceinline__ T epi_rn(T alpha, T acc, T beta, T c) {
#pragma clang fp contract(off)
    const T u = alpha * acc;
    const T v = beta * c;
    return u + v;
}

// Rows are processed CH at a time; each row's chain is j-ascending from 0,
// so chunking changes only which rows share a pass over x, not any row's
// arithmetic.
template <typename T, int CH>
__global__ void gemv_compat_kernel(size_t M, size_t N, size_t bh, int transA, int gemv, T alpha, T beta,
                                   const T* __restrict__ A, const T* __restrict__ x, T* __restrict__ y) {
    const size_t p = blockIdx.y;
    A += p * M * N;
    x += p * N;
    y += p * M;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t rbase = g * bh;
    if (rbase >= M) return;
    for (size_t i0 = 0; i0 < bh; i0 += CH) {
        T acc[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] = 0;
        for (size_t j = 0; j < N; ++j) {
            const T xj = x[j];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const size_t r = rbase + i0 + c;
                const T a = transA ? A[j * M + r] : A[r * N + j];
                acc[c] = mad_rn<T>(acc[c], a, xj);
            }
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const size_t r = rbase + i0 + c;
            if (gemv) {
                y[r] = epi_rn<T>(alpha, acc[c], beta, y[r]);
            } else {
                y[r] = acc[c];
            }
        }
    }
}

int launch_gemv_compat(int dtype, int vl, int transA, size_t M, size_t N, size_t gs, size_t bh, int packed,
                       int gemv, double alpha, double beta, const void* A, const void* x, void* y,
                       hipStream_t st) {
    if (gs == 0 || gs > 1024 || bh == 0 || vl <= 0 || bh % vl != 0 || M % (gs * bh) != 0) {
        set_error("gemv compat: invalid work-group or blocking");
        return ST_INVALID;
    }
    const size_t groups = (M / bh) / gs;
    const int ch = (bh % 8 == 0) ? 8 : (bh % 4 == 0) ? 4 : (bh % 2 == 0) ? 2 : 1;
    dim3 grid((unsigned)groups, (unsigned)packed), block((unsigned)gs);
#define GAT_GEMV_CH(T, C)                                                                              \
    if (ch == C)                                                                                       \
        hipLaunchKernelGGL((gemv_compat_kernel<T, C>), grid, block, 0, st, M, N, bh, transA, gemv, (T)alpha, \
                           (T)beta, (const T*)A, (const T*)x, (T*)y);
    if (dtype == DT_F32) {
        GAT_GEMV_CH(float, 8) else GAT_GEMV_CH(float, 4) else GAT_GEMV_CH(float, 2) else GAT_GEMV_CH(float, 1)
    } else if (dtype == DT_F64) {
        GAT_GEMV_CH(double, 8) else GAT_GEMV_CH(double, 4) else GAT_GEMV_CH(double, 2) else GAT_GEMV_CH(double, 1)
    } else {
        set_error("gemv compat: dtype must be f32 or f64");
        return ST_INVALID;
    }
#undef GAT_GEMV_CH
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
