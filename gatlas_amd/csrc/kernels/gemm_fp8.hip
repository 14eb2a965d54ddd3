// FP8 (OCP e4m3) block quantisation for the FP8 GEMM (gemm_fp8_mfma.hip).
//
// Operands: A8 = op(A) stored M x K row-major, B8 = op(B)^T stored N x K
// row-major (both k-contiguous), scales sA[ceil(M/128)][ceil(K/128)] and
// sB[ceil(N/128)][ceil(K/128)] such that op(A)[i][k] ~ A8[i][k] * sA[i/128][k/128].
//
// quantize_fp8_kernel: one work-group per 128 x 128 block holds the block
// in registers, reduces amax, sets the block scale to the
// smallest power of two >= amax / 448 (1 for an all-zero block) and converts
// x / scale to e4m3 with round-to-nearest-even and saturation at 448,
// writing the k-contiguous layout above.
#include "mfma_common.hpp"

namespace gat {

typedef unsigned char u8;

// FP32 -> OCP e4m3 (bias 7, max 448, subnormal step 2^-9), round to nearest
// even, saturating; NaN -> 0x7F.  Normal results round the FP32 mantissa
// to 3 bits in integer arithmetic (a carry moves into the exponent); below
// 2^-6 the value is a multiple of 2^-9, rint(|x| * 512).
__device__ __forceinline__ u8 f32_to_e4m3(float x) {
    const unsigned ub = __float_as_uint(x);
    const unsigned s = (ub >> 24) & 0x80u;
    const unsigned ab = ub & 0x7FFFFFFFu;
    if (ab > 0x7F800000u) return (u8)(s | 0x7Fu);   // NaN
    if (ab >= 0x43E00000u) return (u8)(s | 0x7Eu);  // |x| >= 448 (and inf)
    if (ab < 0x3C800000u)                           // |x| < 2^-6: 8 -> 0x08 = 2^-6
        return (u8)(s | (unsigned)(int)rintf(__uint_as_float(ab) * 512.0f));
    const unsigned r = (ab + 0x7FFFFu + ((ab >> 20) & 1u)) & ~0xFFFFFu;
    return (u8)(s | ((((r >> 23) - 120u) << 3) | ((r >> 20) & 7u)));  // exponent - 127 + 7
}

// out[r][c] (rows x cols, row-major, ld = cols) from element (r, c) of the
// source at src[r * rs + c * cs]; scales[rb][cb] per 128 x 128 block.
// Work-item t holds rows t/32 + 8i (i < 16) x columns 4*(t%32) .. +3 of
// the block in registers: read straight from a column-contiguous source
// (16 bytes at a time where `vec`), or through LDS from a row-contiguous
// one; each four columns leave as one 4-byte store where aligned.
// COLS: the source is contiguous along columns (cs == 1); the other
// instance stages the block through LDS.
template <bool COLS>
__global__ __launch_bounds__(256) void quantize_fp8_kernel(const float* __restrict__ src, int64_t rs, int64_t cs,
                                                           int rows, int cols, u8* __restrict__ out,
                                                           float* __restrict__ scales, int cblocks, int vec) {
    __shared__ float red[4];
    const int rb = blockIdx.y, cb = blockIdx.x, tid = threadIdx.x;
    const int r0 = rb * 128, c0 = cb * 128;
    const int cq = 4 * (tid & 31), gc = c0 + cq;
    float v[16][4];
    if constexpr (COLS) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int gr = r0 + (tid >> 5) + 8 * i;
            const float* p = src + (int64_t)gr * rs + gc;
            if (vec && gr < rows && gc + 4 <= cols) {
                const f4v a = gload<f4v>(p);
#pragma unroll
                for (int e = 0; e < 4; ++e) v[i][e] = a[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[i][e] = (gr < rows && gc + e < cols) ? gload<float>(p + e) : 0.0f;
            }
        }
    } else {
        __shared__ float tile[128][129];
        if (rs == 1) {
            // 128 columns of 128 contiguous rows: 16-byte loads of four rows
            for (int q = tid; q < 128 * 32; q += 256) {
                const int c = q >> 5, rq = 4 * (q & 31), gr = r0 + rq;
                const float* p = src + (int64_t)(c0 + c) * cs + gr;
                if (vec && c0 + c < cols && gr + 4 <= rows) {
                    const f4v a = gload<f4v>(p);
#pragma unroll
                    for (int e = 0; e < 4; ++e) tile[rq + e][c] = a[e];
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        tile[rq + e][c] = (c0 + c < cols && gr + e < rows) ? gload<float>(p + e) : 0.0f;
                }
            }
        } else {
            for (int q = tid; q < 128 * 128; q += 256) {
                const int r = q & 127, c = q >> 7;
                tile[r][c] = (r0 + r < rows && c0 + c < cols)
                                 ? gload<float>(src + (int64_t)(r0 + r) * rs + (int64_t)(c0 + c) * cs)
                                 : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) v[i][e] = tile[(tid >> 5) + 8 * i][cq + e];
    }
    float amax = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) amax = fmaxf(amax, fabsf(v[i][e]));
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
    const bool wide = (cols % 4) == 0 && ((uintptr_t)out % 4) == 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int gr = r0 + (tid >> 5) + 8 * i;
        if (gr >= rows) continue;
        u8 q[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = f32_to_e4m3(v[i][e] / scale);
        u8* o = out + (int64_t)gr * cols + gc;
        if (wide && gc + 4 <= cols) {
            gstore(o, (unsigned)q[0] | ((unsigned)q[1] << 8) | ((unsigned)q[2] << 16) | ((unsigned)q[3] << 24));
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (gc + e < cols) o[e] = q[e];
        }
    }
}

int launch_quantize_fp8(const float* src, int64_t rs, int64_t cs, int rows, int cols, void* out, float* scales,
                        hipStream_t st) {
    if (rows <= 0 || cols <= 0) return ST_OK;
    const int rbl = (rows + 127) / 128, cbl = (cols + 127) / 128;
    // 16-byte loads along the contiguous direction: aligned base and stride
    const int64_t stride = cs == 1 ? rs : cs;
    const int vec = ((uintptr_t)src % 16 == 0) && (stride % 4 == 0);
    if (cs == 1)
        hipLaunchKernelGGL(quantize_fp8_kernel<true>, dim3(cbl, rbl), dim3(256), 0, st, src, rs, cs, rows, cols,
                           (u8*)out, scales, cbl, vec);
    else
        hipLaunchKernelGGL(quantize_fp8_kernel<false>, dim3(cbl, rbl), dim3(256), 0, st, src, rs, cs, rows, cols,
                           (u8*)out, scales, cbl, vec);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
