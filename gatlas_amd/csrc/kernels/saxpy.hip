// SAXPY / DAXPY:  z = alpha*x + y  (the reference writes a separate Z,
// KernelSaxpyBuffer.hpp:85-92).  Every element is one multiply and one add,
// each rounded (the rounding of the reference's host check,
// KernelSaxpyBuffer.hpp:125-127), so every geometry gives identical bits.
#include "../kernels.hpp"

namespace gat {

// alpha*x + y with the product rounded before the add.  Contraction into a
// fused multiply-add is switched off here: HIP compiles device code with
// fp-contract=fast, which would otherwise fuse the two operations.
template <typename T>
__device__ __forceinline__ T axpy1(T alpha, T x, T y) {
#pragma clang fp contract(off)
    const T p = alpha * x;
    return p + y;
}

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

// Native streaming kernel: each work-item moves UNROLL 16-byte vectors per
// operand, strided by the group size so that every load instruction of a
// wavefront covers 1 KiB of contiguous HBM; all loads are issued before any
// store.  NT selects non-temporal accesses (the streams are touched once).
template <typename T, typename V, int UNROLL, bool NT>
__global__ __launch_bounds__(1024) void axpy_stream_kernel(size_t nvec, T alpha, const V* __restrict__ x,
                                                            const V* __restrict__ y, V* __restrict__ z) {
    constexpr int W = sizeof(V) / sizeof(T);
    const size_t bs = blockDim.x;
    const size_t base = (size_t)blockIdx.x * bs * UNROLL + threadIdx.x;
    V xv[UNROLL], yv[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
        const size_t i = base + (size_t)u * bs;
        if (i < nvec) {
            if (NT) {
                xv[u] = __builtin_nontemporal_load(x + i);
                yv[u] = __builtin_nontemporal_load(y + i);
            } else {
                xv[u] = x[i];
                yv[u] = y[i];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
        const size_t i = base + (size_t)u * bs;
        if (i < nvec) {
            V r;
#pragma unroll
            for (int e = 0; e < W; ++e) r[e] = axpy1<T>(alpha, xv[u][e], yv[u][e]);
            if (NT)
                __builtin_nontemporal_store(r, z + i);
            else
                z[i] = r;
        }
    }
}

template <typename T>
__global__ void axpy_tail_kernel(size_t start, size_t n, T alpha, const T* __restrict__ x,
                                 const T* __restrict__ y, T* __restrict__ z) {
    const size_t i = start + blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) z[i] = axpy1<T>(alpha, x[i], y[i]);
}

bool axpy_cfg_ok(const AxpyCfg& c) {
    return c.wg >= 64 && c.wg <= 1024 && c.wg % 64 == 0 &&
           (c.unroll == 1 || c.unroll == 2 || c.unroll == 4 || c.unroll == 8) &&
           (c.nontemporal == 0 || c.nontemporal == 1);
}

template <typename T, typename V, int U, bool NT>
static void launch_stream(size_t nvec, int wg, T alpha, const T* x, const T* y, T* z, hipStream_t st) {
    const size_t per = (size_t)wg * U;
    const size_t blocks = (nvec + per - 1) / per;
    hipLaunchKernelGGL((axpy_stream_kernel<T, V, U, NT>), dim3((unsigned)blocks), dim3((unsigned)wg), 0, st, nvec,
                       alpha, (const V*)x, (const V*)y, (V*)z);
}

template <typename T, typename V>
static int launch_native_t(const AxpyCfg& c, size_t n, T alpha, const T* x, const T* y, T* z, hipStream_t st) {
    constexpr int W = sizeof(V) / sizeof(T);
    if (!axpy_cfg_ok(c)) {
        set_error("saxpy: unsupported (work-group, unroll, non-temporal)");
        return ST_UNSUPPORTED;
    }
    const bool aligned = ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0) && ((uintptr_t)z % 16 == 0);
    size_t nvec = aligned ? n / W : 0;
    if (nvec) {
#define GAT_AXPY(U, NTV) \
    if (c.unroll == U && c.nontemporal == NTV) launch_stream<T, V, U, (NTV != 0)>(nvec, c.wg, alpha, x, y, z, st);
        GAT_AXPY(1, 0) GAT_AXPY(2, 0) GAT_AXPY(4, 0) GAT_AXPY(8, 0)
        GAT_AXPY(1, 1) GAT_AXPY(2, 1) GAT_AXPY(4, 1) GAT_AXPY(8, 1)
#undef GAT_AXPY
        GAT_HIP_TRY(hipGetLastError());
    }
    const size_t start = nvec * W;
    if (start < n) {
        const size_t rem = n - start;
        hipLaunchKernelGGL((axpy_tail_kernel<T>), dim3((unsigned)((rem + 255) / 256)), dim3(256), 0, st, start,
                           n, alpha, x, y, z);
        GAT_HIP_TRY(hipGetLastError());
    }
    return ST_OK;
}

int launch_saxpy_native(int dtype, size_t n, double alpha, const void* x, const void* y, void* z,
                        hipStream_t st, const AxpyCfg* cfg) {
    if (n == 0) return ST_OK;
    // default: the fastest of the 2^28 sweep (one 16-byte vector per operand
    // per work-item and step, non-temporal)
    const AxpyCfg c = cfg ? *cfg : AxpyCfg{256, 1, 1};
    if (dtype == DT_F32)
        return launch_native_t<float, f4v>(c, n, (float)alpha, (const float*)x, (const float*)y, (float*)z, st);
    if (dtype == DT_F64)
        return launch_native_t<double, d2v>(c, n, alpha, (const double*)x, (const double*)y, (double*)z, st);
    set_error("saxpy: dtype must be f32 or f64");
    return ST_INVALID;
}

// Reference geometry (KernelSaxpyBuffer.hpp:143-225).  Offsets are in
// vectors of VL scalars.
//  1-D (N == VL): work-group gH, work-item `row` of group `g` handles
//      vectors g*bh*gH + row + i*gH, i < bh.
//  2-D: work-group gH x gW, each work-item handles bh x (bw/VL) vectors at
//      row stride gH and column stride gW inside the group's tile.
template <typename T, int VL>
__global__ void saxpy_compat_kernel(size_t M, size_t N, size_t gH, size_t gW, size_t bh, size_t bw,
                                    int twoD, T alpha, const T* __restrict__ X, const T* __restrict__ Y,
                                    T* __restrict__ Z) {
    const size_t p = blockIdx.z;
    const size_t plane = M * N;  // scalars per packed problem
    X += p * plane;
    Y += p * plane;
    Z += p * plane;
    if (!twoD) {
        const size_t row = threadIdx.x, grp = blockIdx.x;
        for (size_t i = 0; i < bh; ++i) {
            const size_t v = grp * bh * gH + row + i * gH;
#pragma unroll
            for (int e = 0; e < VL; ++e) {
                const size_t s = v * VL + e;
                Z[s] = axpy1<T>(alpha, X[s], Y[s]);
            }
        }
    } else {
        const size_t row = threadIdx.y, col = threadIdx.x;
        const size_t brow = blockIdx.y, bcol = blockIdx.x;
        const size_t nv = N / VL;           // vectors per row
        const size_t ww = bw / VL;          // whole vectors per work-item along a row
        for (size_t i = 0; i < bh; ++i)
            for (size_t j = 0; j < ww; ++j) {
                const size_t r = brow * bh * gH + row + i * gH;
                const size_t c = bcol * ww * gW + col + j * gW;
                const size_t v = r * nv + c;
#pragma unroll
                for (int e = 0; e < VL; ++e) {
                    const size_t s = v * VL + e;
                    Z[s] = axpy1<T>(alpha, X[s], Y[s]);
                }
            }
    }
}

template <typename T, int VL>
static int launch_saxpy_compat_t(size_t M, size_t N, size_t gH, size_t gW, size_t bh, size_t bw, int packed,
                                 T alpha, const T* x, const T* y, T* z, hipStream_t st) {
    const bool oneD = (N == (size_t)VL);
    if (oneD) {
        // M vectors of VL scalars; global work-items M/bh, local gH.
        const size_t groups = (M / bh) / gH;
        hipLaunchKernelGGL((saxpy_compat_kernel<T, VL>), dim3((unsigned)groups, 1, (unsigned)packed),
                           dim3((unsigned)gH), 0, st, M, N, gH, gW, bh, bw, 0, alpha, x, y, z);
    } else {
        // global (M/bh, N/bw) work-items, local (gH, gW); x is the column dim.
        const size_t gy = (M / bh) / gH, gx = (N / bw) / gW;
        hipLaunchKernelGGL((saxpy_compat_kernel<T, VL>), dim3((unsigned)gx, (unsigned)gy, (unsigned)packed),
                           dim3((unsigned)gW, (unsigned)gH), 0, st, M, N, gH, gW, bh, bw, 1, alpha, x, y, z);
    }
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_saxpy_compat(int dtype, int vl, size_t M, size_t N, size_t gH, size_t gW, size_t bh, size_t bw,
                        int packed, double alpha, const void* x, const void* y, void* z, hipStream_t st) {
    const bool oneD = (N == (size_t)vl);
    if (gH == 0 || bh == 0 || (!oneD && (gW == 0 || bw == 0)) || gH * (oneD ? 1 : gW) > 1024) {
        set_error("saxpy compat: invalid work-group or blocking");
        return ST_INVALID;
    }
#define GAT_SAXPY_VL(T, V)                                                                             \
    if (vl == V)                                                                                       \
        return launch_saxpy_compat_t<T, V>(M, N, gH, gW, bh, bw, packed, (T)alpha, (const T*)x, (const T*)y, \
                                           (T*)z, st);
    if (dtype == DT_F32) {
        GAT_SAXPY_VL(float, 1)
        GAT_SAXPY_VL(float, 2)
        GAT_SAXPY_VL(float, 4)
    } else if (dtype == DT_F64) {
        GAT_SAXPY_VL(double, 1)
        GAT_SAXPY_VL(double, 2)
        GAT_SAXPY_VL(double, 4)
    }
#undef GAT_SAXPY_VL
    set_error("saxpy compat: unsupported dtype / vector length");
    return ST_INVALID;
}

}  // namespace gat
