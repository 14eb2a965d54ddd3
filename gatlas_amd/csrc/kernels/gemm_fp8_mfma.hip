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
// Interior tiles (whole 128-deep K tiles, 16-byte aligned rows) stage both
// operands straight from HBM into LDS (global_load_lds, rows XOR-swizzled
// as in bf16_dma_tile.hpp) under a software-pipelined K loop; edge tiles
// stage through registers.
#include "bf16_dma_tile.hpp"
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
    static_assert(TM * 32 <= 128 && TN * 32 <= 128 && 128 % (TM * 32) == 0 && 128 % (TN * 32) == 0,
                  "a wave must stay inside one scale block");
    static_assert((BM * (F8_BK / 16)) % NT == 0 && (BN * (F8_BK / 16)) % NT == 0, "16-byte chunks per thread");
};

// Waves per SIMD the LDS footprint allows, capped by 512 registers over the
// accumulators (16 per block) and a working set; the register allocator is
// asked for at least that many.
template <int WM, int WN, int TM, int TN>
constexpr int fp8_waves_per_simd() {
    constexpr int by_lds = (int)(163840 / Fp8Shape<WM, WN, TM, TN>::SMEM) * WM * WN / 4;
    constexpr int by_regs = 512 / (TM * TN * 16 + 128);
    constexpr int w = by_lds < by_regs ? by_lds : by_regs;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// The BM x BN tile at (m0, n0), wholly inside M x N, K a multiple of 128:
// 128-deep stages staged straight into two LDS buffers.  Iteration t runs
// the MFMAs of stage t's first 64-deep half on fragments read earlier while
// reading the second half's; waits for stage t + 1 and for its own reads,
// meets the other waves at a barrier (after which stage t's buffer is free),
// issues the loads of stage t + 2 into that buffer, and runs the second
// half's MFMAs while reading the first fragments of stage t + 1.  A stage's
// loads thus have a whole iteration of matrix work to land.  The MFMA
// sequence is the register-staged kernel's (bit-identical results).
// arb / brb: this wave's scale-block rows.
template <int WM, int WN, int TM, int TN>
__device__ __forceinline__ void fp8_tile_dma(const u8* __restrict__ A, const u8* __restrict__ B,
                                             const float* __restrict__ sA, const float* __restrict__ sB,
                                             float* __restrict__ C, int K, int64_t ldc, float alpha, float beta,
                                             int m0, int n0, int arb, int brb, u8* smem) {
    constexpr int NW = WM * WN, BM = WM * TM * 32, BN = WN * TN * 32;
    typedef DmaOperand<BM, true, NW, u8> OpA;
    typedef DmaOperand<BN, true, NW, u8> OpB;
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    constexpr int LOADS = OpA::PER_WAVE + OpB::PER_WAVE;  // loads per wave per stage
    constexpr int NR = 2 * (TM + TN);                       // LDS reads per 64-deep half
    char* const sb = (char*)smem;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w / WN, wn = w % WN;
    const int KT = K / F8_BK;

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    i32x8 a0[TM], b0[TN], a1[TM], b1[TN];
    // lane l: row l%32 of each block, 32 bytes at k = kk + 32*(l/32)
    auto frags = [&](const char* stage, int kk, i32x8(&a)[TM], i32x8(&b)[TN]) {
        const int c = kk / 16 + 2 * (lane >> 5);
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int row = wm * TM * 32 + i * 32 + (lane & 31);
            const u32x4 lo = OpA::chunk(stage, row, c), hi = OpA::chunk(stage, row, c + 1);
            a[i] = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1], (int)hi[2],
                         (int)hi[3]};
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int row = wn * TN * 32 + j * 32 + (lane & 31);
            const u32x4 lo = OpB::chunk(stage + OpA::BYTES, row, c), hi = OpB::chunk(stage + OpA::BYTES, row, c + 1);
            b[j] = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1], (int)hi[2],
                         (int)hi[3]};
        }
    };
    auto mma = [&](const i32x8(&a)[TM], const i32x8(&b)[TN], int sa, int sbb) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, sa, 0,
                                                                           sbb);
    };
    // the LDS reads (and a stage's loads) spread over the first MFMAs of a
    // half, so the reads land well before the next half needs them
    auto interleave = [&](bool loads) {
        constexpr int M = TM * TN, RPM = (NR + M / 2 - 1) / (M / 2), LPM = (LOADS + M / 2 - 1) / (M / 2);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            if (i < M / 2) {
                __builtin_amdgcn_sched_group_barrier(0x100, RPM, 0);  // LDS reads
                if (loads) __builtin_amdgcn_sched_group_barrier(0x020, LPM, 0);  // global loads into LDS
            }
        }
    };
    auto issue = [&](int t, char* buf) {
        OpA::issue(A, K, m0, t * F8_BK, buf, w, lane);
        OpB::issue(B, K, n0, t * F8_BK, buf + OpA::BYTES, w, lane);
    };
    // Scales: lane l holds those of K tiles 64q + l (chunk q) in a vector
    // register and each iteration takes its own with readlane.  (A scalar
    // load shares its wait counter with the LDS reads and returns out of
    // order, so its use would wait for every read in flight.)
    const float* sAr = sA + (int64_t)arb * KT;
    const float* sBr = sB + (int64_t)brb * KT;
    float sav, sbv;
    auto load_scales = [&](int t0) {
        const int kb = t0 + lane;
        sav = kb < KT ? gload<float>(sAr + kb) : 1.0f;
        sbv = kb < KT ? gload<float>(sBr + kb) : 1.0f;
    };

    load_scales(0);
    issue(0, sb);
    if (KT > 1) {
        issue(1, sb + STAGE);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(LOADS, 0));  // the scales and stage 0 in
    } else {
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));
    }
    __builtin_amdgcn_s_barrier();
    frags(sb, 0, a0, b0);
    for (int t0 = 0; t0 < KT; t0 += 64) {
        if (t0 > 0) {
            load_scales(t0);
            __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 15));
        }
        const int t1 = t0 + 64 < KT ? t0 + 64 : KT;
        for (int t = t0; t < t1; ++t) {
            char* cur = sb + (t & 1) * STAGE;
            const int sa = e8m0_of(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(sav), t & 63)));
            const int sbb = e8m0_of(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(sbv), t & 63)));
            __builtin_amdgcn_sched_barrier(0);
            frags(cur, 64, a1, b1);
            mma(a0, b0, sa, sbb);
            interleave(false);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // stage t + 1 in; this wave's reads of stage t done
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
            // past the last stage the loads repeat it and the reads take
            // what is there, unused; so every iteration has the same shape
            issue(t + 2 < KT ? t + 2 : KT - 1, cur);
            frags(sb + ((t + 1) & 1) * STAGE, 0, a0, b0);
            mma(a1, b1, sa, sbb);
            interleave(true);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the repeated loads land before the wave ends

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * (TM * 32) + i * 32 + mfma32_row(lane, r);
                const int col = n0 + wn * (TN * 32) + j * 32 + mfma32_col(lane);
                float* p = C + (int64_t)row * ldc + col;
                float v = alpha * acc[i][j][r];
                if (beta != 0.0f) v += beta * gload<float>(p);
                gstore(p, v);
            }
}

template <int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(WM* WN * 64)
__attribute__((amdgpu_waves_per_eu(fp8_waves_per_simd<WM, WN, TM, TN>()))) void gemm_fp8_kernel(
    const u8* __restrict__ A, const u8* __restrict__ B, const float* __restrict__ sA, const float* __restrict__ sB,
    float* __restrict__ C, int M, int N, int K, int64_t ldc, float alpha, float beta, int tiles_m, int tiles_n,
    int group_m, int vec_ok, int skip_full) {
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
    if (skip_full && m0 + BM <= M && n0 + BN <= N) return;  // interior tile: the direct-to-LDS kernel's

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
    auto fetch = [&](int t) {
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

    fetch(0);
    sstore(lds8);
    __syncthreads();
    const int arow = wm * (TM * 32) + (lane & 31);
    const int brow = BM + wn * (TN * 32) + (lane & 31);
    const int kh = 32 * (lane >> 5);
    for (int t = 0; t < kblocks; ++t) {
        const u8* buf = lds8 + (t & 1) * ((BM + BN) * F8_LDK);
        if (t + 1 < kblocks) fetch(t + 1);
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

// Interior tiles only (grid = (M/BM) x (N/BN) whole tiles), K a multiple of
// 128, 16-byte aligned rows: direct-to-LDS staging.  A kernel of its own so
// that its registers are allocated for this path alone.
template <int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(WM* WN * 64)
__attribute__((amdgpu_waves_per_eu(fp8_waves_per_simd<WM, WN, TM, TN>()))) void gemm_fp8_dma_kernel(
    const u8* __restrict__ A, const u8* __restrict__ B, const float* __restrict__ sA, const float* __restrict__ sB,
    float* __restrict__ C, int M, int N, int K, int64_t ldc, float alpha, float beta, int tiles_m, int tiles_n,
    int group_m) {
    typedef Fp8Shape<WM, WN, TM, TN> S;
    extern __shared__ __attribute__((aligned(16))) u8 lds8[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave / WN, wn = wave % WN;
    int tmi, tni;
    mfma_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    const int m0 = tmi * S::BM, n0 = tni * S::BN;
    const int arb = (m0 + wm * TM * 32) / 128, brb = (n0 + wn * TN * 32) / 128;
    fp8_tile_dma<WM, WN, TM, TN>(A, B, sA, sB, C, K, ldc, alpha, beta, m0, n0, arb, brb, lds8);
}

// Instantiated configurations (WM, WN, TM, TN); the K depth is always 128.
#define GAT_FP8_CFGS(X) \
    X(2, 2, 2, 2)       \
    X(2, 4, 2, 2)       \
    X(4, 2, 2, 2)       \
    X(2, 4, 4, 2)

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
    // whole tiles through the direct-to-LDS kernel, the rest (or everything,
    // when K is not a multiple of 128 or rows are unaligned) through the
    // register-staged one
    const bool dma = vec_ok && a.K >= F8_BK && a.K % F8_BK == 0;
    const int full_m = (int)(a.M / S::BM), full_n = (int)(a.N / S::BN);
    if (dma && full_m > 0 && full_n > 0) {
        const void* fd = (const void*)gemm_fp8_dma_kernel<WM, WN, TM, TN>;
        if (S::SMEM > 65536) GAT_HIP_TRY(allow_dynamic_lds(fd, S::SMEM));
        const int gm = full_m < 8 ? full_m : 8;
        hipLaunchKernelGGL((gemm_fp8_dma_kernel<WM, WN, TM, TN>), dim3((unsigned)(full_m * full_n)), dim3(S::NT),
                           S::SMEM, st, (const u8*)a.A, (const u8*)a.B, a.scaleA, a.scaleB, (float*)a.C, (int)a.M,
                           (int)a.N, (int)a.K, a.ldc, (float)a.alpha, (float)a.beta, full_m, full_n, gm);
        GAT_HIP_TRY(hipGetLastError());
        if (full_m == tiles_m && full_n == tiles_n) return ST_OK;
    }
    const int skip_full = dma && full_m > 0 && full_n > 0;
    const void* f = (const void*)gemm_fp8_kernel<WM, WN, TM, TN>;
    if (S::SMEM > 65536) GAT_HIP_TRY(allow_dynamic_lds(f, S::SMEM));
    hipLaunchKernelGGL((gemm_fp8_kernel<WM, WN, TM, TN>), dim3((unsigned)(tiles_m * tiles_n)), dim3(S::NT), S::SMEM,
                       st, (const u8*)a.A, (const u8*)a.B, a.scaleA, a.scaleB, (float*)a.C, (int)a.M, (int)a.N,
                       (int)a.K, a.ldc, (float)a.alpha, (float)a.beta, tiles_m, tiles_n, group_m, vec_ok, skip_full);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_fp8(const MfmaCfg& c, const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (c.wk != 1) {
        set_error("k-split wave groups exist for FP32 tiles only");
        return ST_UNSUPPORTED;
    }
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
