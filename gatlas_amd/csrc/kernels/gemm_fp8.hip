// FP8 (OCP e4m3) GEMM with per-128x128-block FP32 scales on gfx950, FP32
// accumulate and output.
//
// Operands: A8 = op(A) stored M x K row-major, B8 = op(B)^T stored N x K
// row-major (both k-contiguous), scales sA[ceil(M/128)][ceil(K/128)] and
// sB[ceil(N/128)][ceil(K/128)] such that op(A)[i][k] ~ A8[i][k] * sA[i/128][k/128].
// The work-group tile is 128 x 128 (2 x 2 wavefronts of 64 x 64, each 2 x 2
// blocks of v_mfma_f32_32x32x16_fp8_fp8) and the K tile is 128 deep, so one
// (sA, sB) pair scales every K tile of a work-group: the tile's products are
// accumulated unscaled in FP32, then added to the running sum times
// sA*sB.  K tiles are double-buffered through LDS (row stride 136 bytes:
// the 16 lanes of a 64-bit LDS read phase fall on distinct banks).
//
// Quantisation (quantize_fp8_kernel): one work-group per 128 x 128 block
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
    if (amax > 0.0f) {
        int e;
        const float m = frexpf(amax / 448.0f, &e);
        scale = ldexpf(1.0f, m == 0.5f ? e - 1 : e);
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

constexpr int F8_BM = 128, F8_BN = 128, F8_BK = 128, F8_LDK = 136;

__global__ __launch_bounds__(256) void gemm_fp8_kernel(const u8* __restrict__ A, const u8* __restrict__ B,
                                                       const float* __restrict__ sA, const float* __restrict__ sB,
                                                       float* __restrict__ C, int M, int N, int K, int64_t ldc,
                                                       float alpha, float beta, int tiles_m, int tiles_n,
                                                       int group_m, int vec_ok) {
    // LDS: 2 buffers x (A 128 rows + B 128 rows) x 136 bytes
    __shared__ __attribute__((aligned(16))) u8 lds[2][2 * 128 * F8_LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int tmi, tni;
    mfma_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    const int m0 = tmi * F8_BM, n0 = tni * F8_BN;
    const int kblocks = (K + F8_BK - 1) / F8_BK;
    const bool full_mn = vec_ok && m0 + F8_BM <= M && n0 + F8_BN <= N;

    f16v acc[2][2], part[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    // each thread moves 4 x 16 bytes of A and of B per K tile
    u32x4 ra[4], rb[4];
    auto gload = [&](int t) {
        const int k0 = t * F8_BK;
        const bool full = full_mn && k0 + F8_BK <= K;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = q * 256 + tid;
            const int r = idx >> 3, kc = (idx & 7) * 16;
            const int gk = k0 + kc;
            if (full) {
                ra[q] = *(const u32x4*)(A + (int64_t)(m0 + r) * K + gk);
                rb[q] = *(const u32x4*)(B + (int64_t)(n0 + r) * K + gk);
            } else {
                u8 ta[16], tb[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    ta[e] = (m0 + r < M && gk + e < K) ? A[(int64_t)(m0 + r) * K + gk + e] : (u8)0;
                    tb[e] = (n0 + r < N && gk + e < K) ? B[(int64_t)(n0 + r) * K + gk + e] : (u8)0;
                }
                ra[q] = *(const u32x4*)ta;
                rb[q] = *(const u32x4*)tb;
            }
        }
    };
    auto sstore = [&](u8* buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = q * 256 + tid;
            const int r = idx >> 3, kc = (idx & 7) * 16;
            u32x2* pa = (u32x2*)(buf + r * F8_LDK + kc);
            u32x2* pb = (u32x2*)(buf + (128 + r) * F8_LDK + kc);
            pa[0] = u32x2{ra[q][0], ra[q][1]};
            pa[1] = u32x2{ra[q][2], ra[q][3]};
            pb[0] = u32x2{rb[q][0], rb[q][1]};
            pb[1] = u32x2{rb[q][2], rb[q][3]};
        }
    };

    gload(0);
    sstore(lds[0]);
    __syncthreads();
    const int arow = wm * 64 + (lane & 31), brow = 128 + wn * 64 + (lane & 31);
    const int kq = 8 * (lane >> 5);
    const int kbl = (K + 127) / 128;
    for (int t = 0; t < kblocks; ++t) {
        const u8* buf = lds[t & 1];
        if (t + 1 < kblocks) gload(t + 1);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) part[i][j][r] = 0.0f;
#pragma unroll
        for (int kk = 0; kk < F8_BK; kk += 16) {
            long a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *(const long*)(buf + (arow + i * 32) * F8_LDK + kk + kq);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *(const long*)(buf + (brow + j * 32) * F8_LDK + kk + kq);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    part[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a[i], b[j], part[i][j], 0, 0, 0);
        }
        const float s = sA[(m0 / 128) * kbl + t] * sB[(n0 / 128) * kbl + t];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][j][r] += s * part[i][j][r];
        if (t + 1 < kblocks) sstore(lds[(t + 1) & 1]);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * 64 + i * 32 + mfma32_row(lane, r);
                const int col = n0 + wn * 64 + j * 32 + mfma32_col(lane);
                if (row < M && col < N) {
                    float* p = C + (int64_t)row * ldc + col;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *p;
                    *p = v;
                }
            }
}

int launch_gemm_fp8(const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) return ST_INVALID;
    if (a.transA || !a.transB || a.lda != a.K || a.ldb != a.K || !a.scaleA || !a.scaleB || a.batch != 1) {
        set_error("fp8 gemm: A must be op(A) as M x K and B op(B) as N x K (k-contiguous), with block scales");
        return ST_INVALID;
    }
    const int tiles_m = (int)((a.M + 127) / 128), tiles_n = (int)((a.N + 127) / 128);
    const int vec_ok = ((uintptr_t)a.A % 16 == 0) && ((uintptr_t)a.B % 16 == 0) && (a.K % 16 == 0);
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    hipLaunchKernelGGL(gemm_fp8_kernel, dim3((unsigned)(tiles_m * tiles_n)), dim3(256), 0, st, (const u8*)a.A,
                       (const u8*)a.B, a.scaleA, a.scaleB, (float*)a.C, (int)a.M, (int)a.N, (int)a.K, a.ldc,
                       (float)a.alpha, (float)a.beta, tiles_m, tiles_n, group_m, vec_ok);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
