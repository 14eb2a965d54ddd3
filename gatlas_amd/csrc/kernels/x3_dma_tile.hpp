// Direct-to-LDS variant of the split-BF16x3 macro tile for interior tiles:
// the FP32 operands go from HBM into LDS unchanged with gfx950's 16-byte
// global_load_lds (the layouts of the FP32 direct-to-LDS tile, 32 k values
// per stage), and each wave splits the FP32 fragments it reads into bf16
// hi/lo right before its matrix instructions.  The register-staged x3 tile
// (bf16_tile.hpp, MODE_X3) instead splits every element once while storing
// a stage to LDS; that costs a register round trip per stage, the packing
// and store instructions, and exposes the stage loads' latency.
//
// The products are the register-staged tile's: the same split (hi =
// bf16(a), lo = bf16(a - hi), both rounded to nearest even), the same lane
// layout of the 32x32x16 operands (lane l holds row l%32 and k values
// kk + 8*(l/32) .. +7 of a 16-deep step) and the same order of the three
// products per 16-deep step (lo*hi, hi*lo for every block, then hi*hi), so
// the two paths agree bit for bit.  A K that is an odd multiple of 16 gets
// a short first stage whose first 16-deep step is all zeros: its products
// add +0 to accumulators that are +0, which changes no bit.
//
// Layouts (sgemm_dma_tile.hpp): k-contiguous operands [row][k], 128 bytes
// per row, chunk c at c ^ ((r >> 1) & 7), a fragment is two 16-byte reads;
// row-contiguous operands [k][row], chunk c of k-row k at c ^ (8 * ((k >> 3)
// & 1)), a fragment is eight 4-byte reads of k-rows kk + 8*(l/32) + e, so
// the two half-waves (k-rows 8 apart) read opposite 128-byte halves of the
// banks.
#pragma once

#include "bf16_tile.hpp"
#include "sgemm_dma_tile.hpp"

namespace gat {

template <int ROWS, bool KCONTIG, int NW>
struct X3DmaOperand;

template <int ROWS, int NW>
struct X3DmaOperand<ROWS, true, NW> : DmaColsF32<ROWS, NW, 1> {};

template <int ROWS, int NW>
struct X3DmaOperand<ROWS, false, NW> : DmaRowsF32<ROWS, NW, 3> {};

template <int WM, int WN, int TM, int TN>
struct X3DmaShape {
    static constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static constexpr size_t SMEM = 2 * (size_t)(BM + BN) * 32 * sizeof(float);
};

// Configurations with a direct-to-LDS x3 routine: both tile sides 64..256
// (whole 1 KiB load blocks per wave) and at most 2 x 2 blocks per wave: the
// fragments of two steps (hi and lo) and 64 accumulators then fit the 256
// registers of two waves per SIMD; 4 x 2 blocks spill.
template <int WM, int WN, int TM, int TN>
constexpr bool x3_has_dma() {
    constexpr int NW = WM * WN, BM = WM * TM * 32, BN = WN * TN * 32;
    return TM * TN <= 4 && BM >= 64 && BN >= 64 && BM <= 256 && BN <= 256 && (BM / 8) % NW == 0 &&
           (BN / 8) % NW == 0;
}

// The eight FP32 values of a fragment split into the bf16 hi and lo operands.
__device__ __forceinline__ void split_frag(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
    u32x4 h, l;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        unsigned hh, ll;
        split_bf16x2(v[2 * e], v[2 * e + 1], hh, ll);
        h[e] = hh;
        l[e] = ll;
    }
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
}

// Computes the BM x BN tile at (m0, n0) of C = alpha*op(A)op(B) + beta*C for
// a tile wholly inside M x N (EDGE: any tile, with M a multiple of 4 when A
// is transposed and N one when B is not -- rows past the matrix are loaded
// as copies of its last rows and not stored), FP32 operands, K >= 32 a
// multiple of 16, 16-byte aligned operand rows; all WM*WN*64 threads of the
// group call it; ends after a barrier (LDS free for the caller's next tile).
template <int WM, int WN, int TM, int TN, bool TA, bool TB, bool EDGE = false>
__device__ __forceinline__ void x3_tile_dma(const float* __restrict__ A, const float* __restrict__ B,
                                            float* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb,
                                            int64_t ldc, float alpha, float beta, int m0, int n0, float* smem) {
    constexpr int NW = WM * WN;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    typedef X3DmaOperand<BM, !TA, NW> OpA;  // op(A) rows are k-contiguous unless A is transposed
    typedef X3DmaOperand<BN, TB, NW> OpB;   // op(B)^T rows are k-contiguous when B is transposed
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    char* const sb = (char*)smem;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w / WN, wn = w % WN;
    (void)M;
    (void)N;

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int kh = K % 32;  // k values of a short first stage (0 or 16)
    const int KT = K / 32 + (kh != 0);
    auto kbase = [&](int st) { return st == 0 ? 0 : st * 32 - (kh ? 32 - kh : 0); };  // first k of stage st

    // Fragments of a 16-deep step go through three forms: FP32 values read
    // from LDS, then bf16 hi/lo operands, then products.  A step's operands
    // are split while the previous step's products run, and its FP32 values
    // are read one step before that, so neither the LDS latency nor the
    // split waits in front of the matrix instructions.
    float ra0[TM][8], rb0[TN][8], ra1[TM][8], rb1[TN][8];
    bf16x8 ah0[TM], al0[TM], bh0[TN], bl0[TN], ah1[TM], al1[TM], bh1[TN], bl1[TN];
    auto frags = [&](const char* stage, int kk, float(&ra)[TM][8], float(&rb)[TN][8]) {
#pragma unroll
        for (int i = 0; i < TM; ++i) OpA::frag8(stage, wm * TM * 32 + i * 32, kk, lane, ra[i]);
#pragma unroll
        for (int j = 0; j < TN; ++j) OpB::frag8(stage + OpA::BYTES, wn * TN * 32 + j * 32, kk, lane, rb[j]);
    };
    auto split = [&](const float(&ra)[TM][8], const float(&rb)[TN][8], bf16x8(&ah)[TM], bf16x8(&al)[TM],
                     bf16x8(&bh)[TN], bf16x8(&bl)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i) split_frag(ra[i], ah[i], al[i]);
#pragma unroll
        for (int j = 0; j < TN; ++j) split_frag(rb[j], bh[j], bl[j]);
    };
    // small terms first, then the leading product (the register path's order)
    auto mma = [&](const bf16x8(&ah)[TM], const bf16x8(&al)[TM], const bf16x8(&bh)[TN], const bf16x8(&bl)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    };
    // one matrix instruction, then a share of the split (and of the LDS
    // reads and stage loads, if any), ...
    constexpr int NMF = 3 * TM * TN;
    constexpr int VPM = (3 * 8 * (TM + TN) + NMF - 1) / NMF;  // split instructions per MFMA
    auto interleave = [&](bool reads_first) {
        // the LDS reads of both steps of the next stage ahead of everything:
        // the first products cover their latency
        if (reads_first) __builtin_amdgcn_sched_group_barrier(0x100, 16 * (TM + TN), 0);
#pragma unroll
        for (int i = 0; i < NMF; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);    // MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, VPM, 0);  // VALU
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);    // global load (to LDS)
        }
    };

    constexpr int LOADS = OpA::PER_WAVE + OpB::PER_WAVE;  // loads per wave per stage
    const float* pa[OpA::PER_WAVE];
    const float* pb[OpB::PER_WAVE];
    OpA::template prep<EDGE>(A, lda, m0, w, lane, pa, M);
    OpB::template prep<EDGE>(B, ldb, n0, w, lane, pb, N);
    auto issue = [&](int st, char* buf) {
        OpA::issue_at(pa, OpA::stage_off(kbase(st), lda), buf, w);
        OpB::issue_at(pb, OpB::stage_off(kbase(st), ldb), buf + OpA::BYTES, w);
    };
    if (kh) {
        OpA::template stage_head<EDGE>(A, lda, m0, 0, kh, sb, threadIdx.x, NW * 64, M);
        OpB::template stage_head<EDGE>(B, ldb, n0, 0, kh, sb + OpA::BYTES, threadIdx.x, NW * 64, N);
    } else {
        issue(0, sb);
    }
    if (KT > 1) {
        issue(1, sb + STAGE);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(LOADS, 0));  // stage 0 in LDS
    } else {
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));
    }
    __builtin_amdgcn_s_barrier();
    frags(sb, 0, ra0, rb0);
    frags(sb, 16, ra1, rb1);
    split(ra0, rb0, ah0, al0, bh0, bl0);

    // Iteration t runs the two 16-deep steps of stage t.  First half: the
    // products of step 0 while step 1's operands are split (its FP32
    // values were read in the previous iteration).  Then each wave waits
    // for stage t+1 and the waves meet at a barrier (every wave has read
    // all of stage t by then); the loads of stage t+2 go into stage t's
    // buffer.  Second half: both steps of stage t+1 are read from LDS, and
    // the products of step 1 run while step 0 of stage t+1 is split.  Past
    // the last stage the loads repeat it and the reads repeat the next
    // buffer (unused), so every iteration has the same shape.
    for (int t = 0; t < KT; ++t) {
        char* cur = sb + (t & 1) * STAGE;
        const char* nxt = sb + ((t + 1) & 1) * STAGE;
        __builtin_amdgcn_sched_barrier(0);
        split(ra1, rb1, ah1, al1, bh1, bl1);
        mma(ah0, al0, bh0, bl0);
        interleave(false);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // stage t+1 in; no read of stage t in flight
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t + 2 < KT ? t + 2 : KT - 1, cur);
        frags(nxt, 0, ra0, rb0);
        frags(nxt, 16, ra1, rb1);
        split(ra0, rb0, ah0, al0, bh0, bl0);
        mma(ah1, al1, bh1, bl1);
        interleave(true);
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the repeated loads land before LDS is reused

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * TM * 32 + i * 32 + mfma32_row(lane, r);
                const int col = n0 + wn * TN * 32 + j * 32 + mfma32_col(lane);
                if (EDGE && (row >= M || col >= N)) continue;
                float* p = C + (int64_t)row * ldc + col;
                float v = alpha * acc[i][j][r];
                if (beta != 0.0f) v += beta * gload<float>(p);
                gstore(p, v);
            }
    __syncthreads();
}

}  // namespace gat
