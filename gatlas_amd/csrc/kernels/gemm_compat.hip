// Reference-geometry GEMM ("compat" engine) for FP32 / FP64.
//
// Geometry of KernelMatmulBuffer.hpp:174-342 / KernelBaseMatmul.cpp:364-376:
//   work-group gs x gs work-items; work-item (row, col) of group
//   (blockRow, blockCol) owns output rows bh*(blockRow*gs + row) + i, i < bh,
//   and columns vl*(blockCol*gs + col) + c, c < vl; the group's macro tile is
//   (gs*bh) x (gs*vl) and K advances gs*vl per LDS stage.
// Arithmetic: every output element is the chain acc = acc + a*b over k
// ascending from acc = 0, each multiply and add separately rounded, then
// C = alpha*acc + beta*C (GEMM form) with the same roundings -- exactly what
// the reference's host check computes (KernelMatmulBuffer.hpp:117-143), so
// the result is bit-identical to it.
// Differences from the reference kernel that do not change any result:
//   * tiles are loaded with coalesced cooperative copies (any work-item may
//     copy any element);
//   * a second barrier after the inner product closes the write-after-read
//     race of the reference's single barrier (KernelMatmulBuffer.hpp:284);
//   * LDS rows carry +1 padding as the reference's localSize() does.
#include "../kernels.hpp"

namespace gat {

// acc + a*b and alpha*acc + beta*c with every product rounded before the
// add.  HIP compiles device code with fp-contract=fast; contraction is
// switched off inside these helpers so no fused multiply-add is formed.
template <typename T>
__device__ __forceinline__ T cmad(T acc, T a, T b) {
#pragma clang fp contract(off)
    const T p = a * b;
    return acc + p;
}
template <typename T>
__device__ __forceinline__ T cepi(T alpha, T acc, T beta, T c) {
#pragma clang fp contract(off)
    const T u = alpha * acc;
    const T v = beta * c;
    return u + v;
}

template <typename T, int VL, int BH>
__global__ __launch_bounds__(256) void gemm_compat_kernel(int M, int N, int K, int gs, int transA, int transB,
                                                          int gemm, T alpha, T beta, const T* __restrict__ A,
                                                          const T* __restrict__ B, T* __restrict__ C) {
    extern __shared__ unsigned char smem_raw[];
    T* smem = (T*)smem_raw;
    const int KS = gs * VL;           // K per stage
    const int TR = gs * BH;           // tile rows
    const int TC = gs * VL;           // tile cols
    const int LA = TR + 1;            // padded LDS row (k-major storage)
    const int LB = TC + 1;
    T* As = smem;                     // [KS][LA]
    T* Bs = smem + KS * LA;           // [KS][LB]

    const int p = blockIdx.z;
    A += (int64_t)p * M * K;
    B += (int64_t)p * K * N;
    C += (int64_t)p * M * N;

    const int row = threadIdx.y, col = threadIdx.x;
    const int tid = row * gs + col, nthr = gs * gs;
    const int r0 = blockIdx.y * TR, c0 = blockIdx.x * TC;

    T acc[BH][VL];
#pragma unroll
    for (int i = 0; i < BH; ++i)
#pragma unroll
        for (int c = 0; c < VL; ++c) acc[i][c] = 0;

    for (int k0 = 0; k0 < K; k0 += KS) {
        // A tile: TR rows x KS k.  Walk the contiguous dimension fastest.
        if (!transA) {
            for (int e = tid; e < TR * KS; e += nthr) {
                const int r = e / KS, k = e % KS;
                As[k * LA + r] = A[(int64_t)(r0 + r) * K + k0 + k];
            }
        } else {
            for (int e = tid; e < TR * KS; e += nthr) {
                const int k = e / TR, r = e % TR;
                As[k * LA + r] = A[(int64_t)(k0 + k) * M + r0 + r];
            }
        }
        // B tile: KS k x TC cols.
        if (!transB) {
            for (int e = tid; e < KS * TC; e += nthr) {
                const int k = e / TC, c = e % TC;
                Bs[k * LB + c] = B[(int64_t)(k0 + k) * N + c0 + c];
            }
        } else {
            for (int e = tid; e < KS * TC; e += nthr) {
                const int c = e / KS, k = e % KS;
                Bs[k * LB + c] = B[(int64_t)(c0 + c) * K + k0 + k];
            }
        }
        __syncthreads();
        for (int k = 0; k < KS; ++k) {
            T a[BH], b[VL];
#pragma unroll
            for (int i = 0; i < BH; ++i) a[i] = As[k * LA + row * BH + i];
#pragma unroll
            for (int c = 0; c < VL; ++c) b[c] = Bs[k * LB + col * VL + c];
#pragma unroll
            for (int i = 0; i < BH; ++i)
#pragma unroll
                for (int c = 0; c < VL; ++c) acc[i][c] = cmad<T>(acc[i][c], a[i], b[c]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < BH; ++i)
#pragma unroll
        for (int c = 0; c < VL; ++c) {
            T* out = C + (int64_t)(r0 + row * BH + i) * N + c0 + col * VL + c;
            *out = gemm ? cepi<T>(alpha, acc[i][c], beta, *out) : acc[i][c];
        }
}

#define GAT_COMPAT_BH(X, T, VL) \
    X(T, VL, 1) X(T, VL, 2) X(T, VL, 3) X(T, VL, 4) X(T, VL, 5) X(T, VL, 6) X(T, VL, 7) X(T, VL, 8) \
    X(T, VL, 9) X(T, VL, 10) X(T, VL, 11) X(T, VL, 12)

bool compat_has(int dtype, const CompatCfg& c) {
    return (dtype == DT_F32 || dtype == DT_F64) && (c.vl == 1 || c.vl == 2 || c.vl == 4) && c.bh >= 1 &&
           c.bh <= 12 && c.gs >= 1 && c.gs <= 16;
}

template <typename T, int VL, int BH>
static int launch_compat_t(const CompatCfg& cfg, const GemmArgs& a, hipStream_t st) {
    const int gs = cfg.gs;
    const size_t smem = ((size_t)gs * VL * (gs * BH + 1) + (size_t)gs * VL * (gs * VL + 1)) * sizeof(T);
    if (smem > 160 * 1024) {
        set_error("gemm compat: LDS tile exceeds 160 KiB");
        return ST_INVALID;
    }
    auto f = gemm_compat_kernel<T, VL, BH>;
    if (smem > 65536) GAT_HIP_TRY(allow_dynamic_lds((const void*)f, smem));
    dim3 grid((unsigned)(a.N / (gs * VL)), (unsigned)(a.M / (gs * BH)), (unsigned)a.batch);
    dim3 block((unsigned)gs, (unsigned)gs);
    const int gemm = !(a.alpha == 1.0 && a.beta == 0.0);
    hipLaunchKernelGGL(f, grid, block, smem, st, (int)a.M, (int)a.N, (int)a.K, gs, a.transA, a.transB, gemm,
                       (T)a.alpha, (T)a.beta, (const T*)a.A, (const T*)a.B, (T*)a.C);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_compat(int dtype, const CompatCfg& cfg, const GemmArgs& a, hipStream_t st) {
    if (!compat_has(dtype, cfg)) {
        set_error("gemm compat: unsupported (dtype, gs, bh, vl)");
        return ST_UNSUPPORTED;
    }
    const int64_t TR = (int64_t)cfg.gs * cfg.bh, TC = (int64_t)cfg.gs * cfg.vl;
    // KernelBaseMatmul.cpp:263-291 divisibility; leading dims are the packed ones.
    if (a.M % TR || a.N % TC || a.K % TC || a.M < TR || a.N < TC || a.K < TC ||
        a.lda != (a.transA ? a.M : a.K) || a.ldb != (a.transB ? a.K : a.N) || a.ldc != a.N) {
        set_error("gemm compat: shape does not fit the reference blocking");
        return ST_INVALID;
    }
    if (a.M == 0 || a.N == 0 || a.batch <= 0) return ST_OK;
#define X(T, V, H) \
    if (cfg.vl == V && cfg.bh == H) return launch_compat_t<T, V, H>(cfg, a, st);
    if (dtype == DT_F32) {
        GAT_COMPAT_BH(X, float, 1)
        GAT_COMPAT_BH(X, float, 2)
        GAT_COMPAT_BH(X, float, 4)
    } else {
        GAT_COMPAT_BH(X, double, 1)
        GAT_COMPAT_BH(X, double, 2)
        GAT_COMPAT_BH(X, double, 4)
    }
#undef X
    return ST_UNSUPPORTED;
}

}  // namespace gat
