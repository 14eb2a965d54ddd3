// FP8 (OCP e4m3) GEMM with per-128x128-block FP32 scales on gfx950, FP32
// accumulate and output: C = alpha * (A8 .* sA) (B8 .* sB)^T + beta * C with
// A8 = op(A) as M x K and B8 = op(B)^T as N x K, both k-contiguous bytes
// (quantised by gemm_fp8.hip).
//
// Work-group: WM x WN wavefronts, each wave TM x TN blocks of
// v_mfma_scale_f32_32x32x64_f8f6f4 (both operands e4m3), macro tile
// (WM*TM*32) x (WN*TN*32).  A wave spans at most 128 rows and 128 columns
// inside one 128-aligned block, and the K tile is 128 deep = one scale
// block, so a wave meets a single (sA, sB) pair per K tile.  The pair goes
// into the instruction's E8M0 scale operands -- each scale is a power of two,
// so its E8M0 byte is its FP32 exponent field -- and the products reach the
// accumulators already scaled.
// Operand layout: lane l holds row l%32 of its A block and column l%32 of
// its B block, 32 bytes each at k = kk + 32*(l/32) .. +31; A and B use the
// same k assignment, so the 64-deep sum pairs bytes of equal k.  K tiles are
// double-buffered through LDS, rows k-contiguous at a 144-byte stride.
#include "mfma_common.hpp"

namespace gat {

typedef unsigned char u8;
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int F8_BK = 128, F8_LDK = 144;

__device__ __forceinline__ int e8m0_of(float pow2) { return (int)((__float_as_uint(pow2) >> 23) & 0xFF); }

template <int WM, int WN, int TM, int TN>
struct Fp8Shape {
    static constexpr int NT = WM * WN * 64;
    static constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static constexpr size_t SMEM = 2 * (size_t)(BM + BN) * F8_LDK;
    static_assert(TM * 32 <= 128 && TN * 32 <= 128, "a wave must stay inside one scale block");
    static_assert((BM * (F8_BK / 16)) % NT == 0 && (BN * (F8_BK / 16)) % NT == 0, "16-byte chunks per thread");
};

template <int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(WM* WN * 64) void gemm_fp8_kernel(const u8* __restrict__ A, const u8* __restrict__ B,
                                                               const float* __restrict__ sA,
                                                               const float* __restrict__ sB, float* __restrict__ C,
                                                               int M, int N, int K, int64_t ldc, float alpha,
                                                               float beta, int tiles_m, int tiles_n, int group_m,
                                                               int vec_ok) {
    typedef Fp8Shape<WM, WN, TM, TN> S;
    constexpr int NT = S::NT, BM = S::BM, BN = S::BN;
    constexpr int QA = BM * (F8_BK / 16) / NT, QB = BN * (F8_BK / 16) / NT;  // chunks per thread
    extern __shared__ __attribute__((aligned(16))) u8 lds8[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    int tmi, tni;
    mfma_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    const int m0 = tmi * BM, n0 = tni * BN;
    const int kblocks = (K + F8_BK - 1) / F8_BK;
    const bool full_mn = vec_ok && m0 + BM <= M && n0 + BN <= N;
    // this wave's scale-block row of A and of B (clamped: rows past the
    // edge are computed but never stored)
    const int mbl = (M + 127) / 128, nbl = (N + 127) / 128;
    const int arb = min((m0 + wm * TM * 32) / 128, mbl - 1), brb = min((n0 + wn * TN * 32) / 128, nbl - 1);

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    u32x4 ra[QA], rb[QB];
    auto chunk = [&](const u8* __restrict__ src, int row, int rlim, int gk, bool full) -> u32x4 {
        const u8* p = src + (int64_t)row * K + gk;
        if (full) return *(const u32x4*)p;
        if (vec_ok && row < rlim && gk + 16 <= K) return *(const u32x4*)p;
        u32x4 v = {0u, 0u, 0u, 0u};
        if (row < rlim)
            for (int e = 0; e < 16 && gk + e < K; ++e) v[e >> 2] |= (unsigned)p[e] << (8 * (e & 3));
        return v;
    };
    auto gload = [&](int t) {
        const int k0 = t * F8_BK;
        const bool full = full_mn && k0 + F8_BK <= K;
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int idx = q * NT + tid;
            ra[q] = chunk(A, m0 + (idx >> 3), M, k0 + (idx & 7) * 16, full);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int idx = q * NT + tid;
            rb[q] = chunk(B, n0 + (idx >> 3), N, k0 + (idx & 7) * 16, full);
        }
    };
    auto sstore = [&](u8* buf) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int idx = q * NT + tid;
            *(u32x4*)(buf + (idx >> 3) * F8_LDK + (idx & 7) * 16) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int idx = q * NT + tid;
            *(u32x4*)(buf + (BM + (idx >> 3)) * F8_LDK + (idx & 7) * 16) = rb[q];
        }
    };

    gload(0);
    sstore(lds8);
    __syncthreads();
    const int arow = wm * (TM * 32) + (lane & 31);
    const int brow = BM + wn * (TN * 32) + (lane & 31);
    const int kh = 32 * (lane >> 5);
    for (int t = 0; t < kblocks; ++t) {
        const u8* buf = lds8 + (t & 1) * ((BM + BN) * F8_LDK);
        if (t + 1 < kblocks) gload(t + 1);
        const int sa = e8m0_of(sA[(int64_t)arb * kblocks + t]);
        const int sb = e8m0_of(sB[(int64_t)brb * kblocks + t]);
#pragma unroll
        for (int kk = 0; kk < F8_BK; kk += 64) {
            i32x8 b[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const u8* p = buf + (brow + j * 32) * F8_LDK + kk + kh;
                const u32x4 lo = *(const u32x4*)p, hi = *(const u32x4*)(p + 16);
                b[j] = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1], (int)hi[2],
                             (int)hi[3]};
            }
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const u8* p = buf + (arow + i * 32) * F8_LDK + kk + kh;
                const u32x4 lo = *(const u32x4*)p, hi = *(const u32x4*)(p + 16);
                const i32x8 a = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1],
                                      (int)hi[2], (int)hi[3]};
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b[j], acc[i][j], 0, 0, 0, sa, 0,
                                                                               sb);
            }
        }
        if (t + 1 < kblocks) sstore(lds8 + ((t + 1) & 1) * ((BM + BN) * F8_LDK));
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * (TM * 32) + i * 32 + mfma32_row(lane, r);
                const int col = n0 + wn * (TN * 32) + j * 32 + mfma32_col(lane);
                if (row < M && col < N) {
                    float* p = C + (int64_t)row * ldc + col;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *p;
                    *p = v;
                }
            }
}

// Instantiated configurations (WM, WN, TM, TN); the K depth is always 128.
#define GAT_FP8_CFGS(X) \
    X(2, 2, 2, 2)       \
    X(2, 4, 2, 2)       \
    X(4, 2, 2, 2)

static const MfmaCfg kFp8Cfgs[] = {
#define X(wm, wn, tm, tn) {wm, wn, tm, tn, F8_BK, 32},
    GAT_FP8_CFGS(X)
#undef X
};

int mfma_fp8_num_cfgs() { return (int)(sizeof(kFp8Cfgs) / sizeof(kFp8Cfgs[0])); }
MfmaCfg mfma_fp8_cfg(int i) { return kFp8Cfgs[i]; }
bool mfma_fp8_has(const MfmaCfg& c) {
    for (const MfmaCfg& k : kFp8Cfgs)
        if (k.waves_m == c.waves_m && k.waves_n == c.waves_n && k.tm == c.tm && k.tn == c.tn && k.bk == c.bk &&
            k.blk == c.blk)
            return true;
    return false;
}

template <int WM, int WN, int TM, int TN>
static int launch_fp8_one(const GemmArgs& a, hipStream_t st) {
    typedef Fp8Shape<WM, WN, TM, TN> S;
    const int tiles_m = (int)((a.M + S::BM - 1) / S::BM), tiles_n = (int)((a.N + S::BN - 1) / S::BN);
    const int vec_ok = ((uintptr_t)a.A % 16 == 0) && ((uintptr_t)a.B % 16 == 0) && (a.K % 16 == 0);
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    const void* f = (const void*)gemm_fp8_kernel<WM, WN, TM, TN>;
    if (S::SMEM > 65536) GAT_HIP_TRY(allow_dynamic_lds(f, S::SMEM));
    hipLaunchKernelGGL((gemm_fp8_kernel<WM, WN, TM, TN>), dim3((unsigned)(tiles_m * tiles_n)), dim3(S::NT), S::SMEM,
                       st, (const u8*)a.A, (const u8*)a.B, a.scaleA, a.scaleB, (float*)a.C, (int)a.M, (int)a.N,
                       (int)a.K, a.ldc, (float)a.alpha, (float)a.beta, tiles_m, tiles_n, group_m, vec_ok);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_fp8(const MfmaCfg& c, const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) return ST_INVALID;
    if (a.transA || !a.transB || a.lda != a.K || a.ldb != a.K || !a.scaleA || !a.scaleB || a.batch != 1) {
        set_error("fp8 gemm: A must be op(A) as M x K and B op(B) as N x K (k-contiguous), with block scales");
        return ST_INVALID;
    }
#define X(p_wm, p_wn, p_tm, p_tn)                                                                       \
    if (c.waves_m == p_wm && c.waves_n == p_wn && c.tm == p_tm && c.tn == p_tn && c.bk == F8_BK &&    \
        c.blk == 32)                                                                                  \
        return launch_fp8_one<p_wm, p_wn, p_tm, p_tn>(a, st);
    GAT_FP8_CFGS(X)
#undef X
    set_error("fp8 gemm: configuration not instantiated");
    return ST_UNSUPPORTED;
}

}  // namespace gat
