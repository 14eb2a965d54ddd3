// Device-side synthetic input generation and precision conversions.
//
// fill_uniform: element idx of stream `stream` under `seed` is
//   z = splitmix64_finalize(seed*G ^ (stream << 48) ^ idx + G)
//   u = (z >> 40) * 2^-24            (24-bit uniform in [0,1), exact in FP32)
//   v = lo + (hi - lo) * u
// For [lo,hi] = [-1,1] the value 2u-1 is exact in FP32, so a host restatement
// reproduces every element bit-for-bit.
#include "../kernels.hpp"

namespace gat {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void fill_uniform_kernel(T* __restrict__ out, size_t n, uint64_t seed, uint64_t stream,
                                    double lo, double hi) {
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
    const uint64_t key = (seed * G) ^ (stream << 48);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t z = mix64((key ^ (uint64_t)i) + G);
        const double u = (double)(z >> 40) * (1.0 / 16777216.0);
        out[i] = (T)(lo + (hi - lo) * u);
    }
}

// Box-Muller normal N(0,1) from two uniforms of the same hash family
// (sub-streams 2*idx and 2*idx+1), computed in FP64 then rounded.
template <typename T>
__global__ void fill_normal_kernel(T* __restrict__ out, size_t n, uint64_t seed, uint64_t stream) {
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
    const uint64_t key = (seed * G) ^ (stream << 48) ^ (1ULL << 47);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t z1 = mix64((key ^ (uint64_t)(2 * i)) + G);
        const uint64_t z2 = mix64((key ^ (uint64_t)(2 * i + 1)) + G);
        const double u1 = ((double)(z1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0,1]
        const double u2 = (double)(z2 >> 11) * (1.0 / 9007199254740992.0);
        out[i] = (T)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
}

// Scales column c of a rows x cols row-major matrix by 10^s_c with
// s_c uniform in (-3, 3) from stream `stream` (one draw per column).
__global__ void scale_columns_kernel(float* __restrict__ a, size_t rows, size_t cols,
                                     uint64_t seed, uint64_t stream) {
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
    const uint64_t key = (seed * G) ^ (stream << 48);
    const size_t n = rows * cols;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t c = i % cols;
        const uint64_t z = mix64((key ^ (uint64_t)c) + G);
        const double u = (double)(z >> 40) * (1.0 / 16777216.0);
        a[i] = (float)((double)a[i] * pow(10.0, -3.0 + 6.0 * u));
    }
}

static inline dim3 grid_for(size_t n) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    return dim3((unsigned)blocks);
}

int launch_fill_uniform(int dtype, void* out, size_t n, uint64_t seed, uint64_t stream, double lo,
                        double hi, hipStream_t st) {
    if (n == 0) return ST_OK;
    if (dtype == DT_F32)
        hipLaunchKernelGGL(fill_uniform_kernel<float>, grid_for(n), dim3(256), 0, st, (float*)out, n,
                           seed, stream, lo, hi);
    else if (dtype == DT_F64)
        hipLaunchKernelGGL(fill_uniform_kernel<double>, grid_for(n), dim3(256), 0, st, (double*)out, n,
                           seed, stream, lo, hi);
    else
        return ST_INVALID;
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_fill_normal(int dtype, void* out, size_t n, uint64_t seed, uint64_t stream, hipStream_t st) {
    if (n == 0) return ST_OK;
    if (dtype == DT_F32)
        hipLaunchKernelGGL(fill_normal_kernel<float>, grid_for(n), dim3(256), 0, st, (float*)out, n, seed,
                           stream);
    else if (dtype == DT_F64)
        hipLaunchKernelGGL(fill_normal_kernel<double>, grid_for(n), dim3(256), 0, st, (double*)out, n,
                           seed, stream);
    else
        return ST_INVALID;
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_scale_columns(float* a, size_t rows, size_t cols, uint64_t seed, uint64_t stream,
                         hipStream_t st) {
    const size_t n = rows * cols;
    if (n == 0) return ST_OK;
    hipLaunchKernelGGL(scale_columns_kernel, grid_for(n), dim3(256), 0, st, a, rows, cols, seed, stream);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
