// FP8 (OCP e4m3) block quantisation for the FP8 GEMM (gemm_fp8_mfma.hip).
//
// Operands: A8 = op(A) stored M x K row-major, B8 = op(B)^T stored N x K
// row-major (both k-contiguous), scales sA[ceil(M/128)][ceil(K/128)] and
// sB[ceil(N/128)][ceil(K/128)] such that op(A)[i][k] ~ A8[i][k] * sA[i/128][k/128].
//
// quantize_fp8_kernel: one work-group per 128 x 128 block
// stages the block in LDS, reduces amax, sets the block scale to the
// smallest power of two >= amax / 448 (1 for an all-zero block) and converts
// x / scale to e4m3 with round-to-nearest-even and saturation at 448,
// writing the k-contiguous layout above.
#include "mfma_common.hpp"

namespace gat {

typedef unsigned char u8;

// FP32 -> OCP e4m3 (bias 7, max 448, subnormal step 2^-9), round to nearest
// even, saturating; NaN -> 0x7F.
__device__ __forceinline__ u8 f32_to_e4m3(float x) {
    const u8 s = (u8)((__float_as_uint(x) >> 24) & 0x80);
    const float a = fabsf(x);
    if (!(a == a)) return (u8)(s | 0x7F);
    if (a >= 448.0f) return (u8)(s | 0x7E);
    if (a < 0.015625f) return (u8)(s | (u8)rintf(a * 512.0f));  // 8 -> 0x08 = 2^-6
    int e;
    frexpf(a, &e);  // a in [2^(e-1), 2^e)
    int E = e - 1;
    float q = rintf(ldexpf(a, 3 - E));  // [8, 16]
    if (q >= 16.0f) {
        q = 8.0f;
        E += 1;
    }
    if (E > 8) return (u8)(s | 0x7E);
    int b = ((E + 7) << 3) | ((int)q - 8);
    if (b > 0x7E) b = 0x7E;
    return (u8)(s | b);
}

// out[r][c] (rows x cols, row-major, ld = cols) from element (r, c) of the
// source at src[r * rs + c * cs]; scales[rb][cb] per 128 x 128 block.
__global__ __launch_bounds__(256) void quantize_fp8_kernel(const float* __restrict__ src, int64_t rs, int64_t cs,
                                                           int rows, int cols, u8* __restrict__ out,
                                                           float* __restrict__ scales, int cblocks) {
    __shared__ float tile[128][129];
    __shared__ float red[4];
    const int rb = blockIdx.y, cb = blockIdx.x, tid = threadIdx.x;
    const int r0 = rb * 128, c0 = cb * 128;
    const bool col_fast = (cs == 1);  // source contiguous along c
    float amax = 0.0f;
    for (int e = tid; e < 128 * 128; e += 256) {
        const int r = col_fast ? e / 128 : e % 128;
        const int c = col_fast ? e % 128 : e / 128;
        float v = 0.0f;
        if (r0 + r < rows && c0 + c < cols) v = src[(int64_t)(r0 + r) * rs + (int64_t)(c0 + c) * cs];
        tile[r][c] = v;
        amax = fmaxf(amax, fabsf(v));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = amax;
    __syncthreads();
    amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    // smallest power of two >= amax/448: scaling is then exact in both
    // directions, and |x / scale| <= 448 up to the rounding of amax/448
    float scale = 1.0f;
    if (amax > 0.0f) {  // kept normal (>= 2^-126): an E8M0 byte of its own exponent field
        int e;
        const float m = frexpf(amax / 448.0f, &e);
        scale = ldexpf(1.0f, m == 0.5f ? e - 1 : e);
        scale = fmaxf(scale, 0x1p-126f);
    }
    if (tid == 0) scales[rb * cblocks + cb] = scale;
    for (int e = tid; e < 128 * 128; e += 256) {
        const int r = e / 128, c = e % 128;
        if (r0 + r < rows && c0 + c < cols) out[(int64_t)(r0 + r) * cols + c0 + c] = f32_to_e4m3(tile[r][c] / scale);
    }
}

int launch_quantize_fp8(const float* src, int64_t rs, int64_t cs, int rows, int cols, void* out, float* scales,
                        hipStream_t st) {
    if (rows <= 0 || cols <= 0) return ST_OK;
    const int rbl = (rows + 127) / 128, cbl = (cols + 127) / 128;
    hipLaunchKernelGGL(quantize_fp8_kernel, dim3(cbl, rbl), dim3(256), 0, st, src, rs, cs, rows, cols, (u8*)out,
                       scales, cbl);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
