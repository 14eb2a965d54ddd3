// Direct-to-LDS variant of the BF16 macro tile (BF16 inputs, BK = 64) for
// interior tiles: operands go from HBM into LDS with gfx950's 16-byte
// global_load_lds (no register staging, no LDS store instructions), so the
// next K stage lands in LDS while the MFMAs run on the current one.
//
// A global_load_lds of a wavefront writes 1 KiB contiguous in LDS, lane l's
// 16 bytes at base + 16*l; the layout is chosen by which 16-byte chunk of
// global memory each lane fetches.
//   k-contiguous operand (A not transposed, B transposed): [row][k], 128
//     bytes per row (8 chunks), chunk c of row r stored at position
//     c ^ (r & 7): the 32 rows x 2 chunks of one fragment read cover every
//     bank group evenly.
//   row-contiguous operand (A transposed, B not transposed): [k][row] as in
//     memory, chunk c of k-row k stored at c ^ (4 * (k & 3)) (mod chunks
//     per k-row); fragments are
//     read with ds_read_b64_tr_b16.  Within each 16-lane group, lane i's
//     half s is element (i % 4) of the 8 bytes addressed by lane 4s + i/4,
//     so lane j addressing (k-row kb + j/4, row n0 + 4*(j % 4)) hands lane
//     i rows n0 + i, k-rows kb .. kb+3; two reads make the 8 k values of
//     an MFMA operand.
// The MFMA operand layout, accumulators and epilogue are bf16_tile's.
#pragma once

#include "bf16_tile.hpp"

namespace gat {

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4_t;

// One operand of a K stage: ROWS rows of RB bytes of k (RB = 128: 64 bf16
// or 128 fp8 values; RB = 64: 32 bf16 values), loaded by NW wavefronts; T
// is the element type.  The row-contiguous layout exists for 2-byte
// elements only, and holds RB / 2 k-rows.
//   k-contiguous, RB = 128: chunk c of row r at position c ^ (r & 7).
//   k-contiguous, RB = 64: chunk c of row r at position c ^ ((r >> 2) & 3):
//     a 256-byte pass over the banks holds four rows, and the 16 rows that
//     16 lanes of a fragment read fall on 16 different 16-byte bank groups.
template <int ROWS, bool KCONTIG, int NW, typename T = __bf16, int RB = 128>
struct DmaOperand {
    static_assert(RB == 128 || (RB == 64 && sizeof(T) == 2), "128-byte rows, or 64-byte rows of bf16");
    static constexpr int EPC = 16 / (int)sizeof(T);  // elements per 16-byte chunk
    static constexpr int CH = RB / 16;               // chunks per row, k-contiguous layout
    static constexpr int RPB = 1024 / RB;            // rows per 1 KiB block, k-contiguous layout
    static constexpr int KD = RB / (int)sizeof(T);   // k values per stage
    static constexpr int BYTES = ROWS * RB;
    static constexpr int BLOCKS = BYTES / 1024;  // one wavefront instruction each
    static constexpr int PER_WAVE = BLOCKS / NW;
    static_assert(PER_WAVE * NW == BLOCKS, "blocks must split evenly over the wavefronts");
    static constexpr int CPR = ROWS / 8;  // chunks per k-row, row-contiguous layout
    static constexpr int KPB = 64 / CPR;  // k-rows per block, row-contiguous layout
    static_assert(KCONTIG || (sizeof(T) == 2 && CPR >= 8 && CPR <= 64 && 64 % CPR == 0),
                  "row-contiguous tile: 2-byte elements, 64..512 rows");

    // position swizzle of row r, k-contiguous layout
    __device__ static __forceinline__ int ksw(int r) { return RB == 128 ? (r & 7) : ((r >> 2) & 3); }

    // A k-row of the row-contiguous layout is ROWS*2 bytes, a multiple of
    // the 256 bytes of one pass over the 64 banks, so every k-row starts at
    // bank 0.  A fragment read touches 64 contiguous bytes of each of k-rows
    // kb..kb+3 and kb+8..kb+11; XOR-ing their chunks by 4*(k mod 4) puts the
    // four k-rows of each 16-lane group on four different 64-byte bank
    // windows, so the 512 bytes of one read take two passes, not four.
    __device__ static __forceinline__ int tswz(int k) { return (4 * (k & 3)) & (CPR - 1); }

    // Where a tile reaches past the last row (EDGE, rlim rows in all): a
    // k-contiguous operand reads its last row there, a row-contiguous one
    // (rlim a multiple of EPC) its last whole chunk, so every load stays
    // inside the matrix and the rows past it hold copies.
    template <bool EDGE>
    __device__ static __forceinline__ int srow(int r, int rlim) {
        return EDGE ? (r < rlim ? r : rlim - 1) : r;
    }
    template <bool EDGE>
    __device__ static __forceinline__ int chunk_row(int r, int rlim) {
        return EDGE ? (r < rlim - EPC ? r : rlim - EPC) : r;
    }

    // Wavefront w fetches its blocks of the stage starting at k0 into `lds`.
    template <bool EDGE = false>
    __device__ static __forceinline__ void issue(const T* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane, int rlim = 0) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int b = i * NW + w;
            const T* g;
            if constexpr (KCONTIG) {
                const int row = RPB * b + lane / CH;
                g = src + (int64_t)srow<EDGE>(r0 + row, rlim) * ld + k0 + EPC * ((lane % CH) ^ ksw(row));
            } else {
                const int k = b * KPB + lane / CPR;
                g = src + (int64_t)(k0 + k) * ld + chunk_row<EDGE>(r0 + EPC * ((lane % CPR) ^ tswz(k)), rlim);
            }
            __builtin_amdgcn_global_load_lds((glb_void_t*)g, (lds_void_t*)(lds + b * 1024), 16, 0, 0);
        }
    }

    // The stage of k values [k0, k0 + klen) (klen < KD a multiple of EPC;
    // the rest of the KD zero) written into `lds` in this operand's layout
    // through registers, by all nt work-items: the first stage of a K range
    // that is not a multiple of KD.
    template <bool EDGE = false>
    __device__ static __forceinline__ void stage_head(const T* src, int64_t ld, int r0, int k0, int klen, char* lds,
                                                      int tid, int nt, int rlim = 0) {
        if constexpr (KCONTIG) {
            for (int q = tid; q < ROWS * CH; q += nt) {
                const int row = q / CH, c = q % CH;
                u32x4 v = {0u, 0u, 0u, 0u};
                if (EPC * c < klen)
                    v = gload<u32x4>(src + (int64_t)srow<EDGE>(r0 + row, rlim) * ld + k0 + EPC * c);
                *(u32x4*)(lds + row * RB + 16 * (c ^ ksw(row))) = v;
            }
        } else {
            for (int q = tid; q < KD * CPR; q += nt) {
                const int k = q / CPR, c = q % CPR;
                u32x4 v = {0u, 0u, 0u, 0u};
                if (k < klen) v = gload<u32x4>(src + (int64_t)(k0 + k) * ld + chunk_row<EDGE>(r0 + EPC * c, rlim));
                *(u32x4*)(lds + k * (ROWS * 2) + 16 * (c ^ tswz(k))) = v;
            }
        }
    }

    // 16-byte chunk c (k bytes 16c .. 16c+15) of row `row` (k-contiguous layout)
    __device__ static __forceinline__ u32x4 chunk(const char* lds, int row, int c) {
        return *(const u32x4*)(lds + row * RB + 16 * (c ^ ksw(row)));
    }

    // BF16 MFMA operand (8 k values of row rbase + lane%32 at k = kk + 8*(lane/32))
    __device__ static __forceinline__ bf16x8 frag(const char* lds, int rbase, int kk, int lane) {
        if constexpr (KCONTIG) {
            const int row = rbase + (lane & 31);  // rbase a multiple of 32
            const int c = (kk >> 3) + (lane >> 5);
            return *(const bf16x8*)(lds + row * RB + 16 * (c ^ ksw(lane & 31)));
        } else {
            const int g = lane >> 4, j = lane & 15;
            const int n = rbase + 16 * (g & 1) + 4 * (j & 3);
            const int k = kk + 8 * (g >> 1) + (j >> 2);
            const int off0 = k * (ROWS * 2) + 16 * ((n >> 3) ^ tswz(k)) + 2 * (n & 7);
            const int off1 = off0 + 4 * (ROWS * 2);  // k + 4: same k mod 4, same swizzle
            const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4_t*)(lds + off0));
            const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4_t*)(lds + off1));
            const s16x4 v[2] = {lo, hi};
            return *(const bf16x8*)v;
        }
    }
};

template <int WM, int WN, int TM, int TN>
struct Bf16DmaShape {
    static constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static constexpr size_t SMEM = 2 * (size_t)(BM + BN) * 64 * sizeof(__bf16);
};

// Computes the BM x BN tile at (m0, n0) of C = alpha*op(A)op(B) + beta*C for
// a tile wholly inside M x N (EDGE: any tile, with M a multiple of 8 when A
// is transposed and N one when B is not -- rows past the matrix are loaded
// as copies of its last rows and not stored), K >= 64 a multiple of 8,
// 16-byte aligned rows; all WM*WN*64 threads of the group call it; each wave returns with none
// of its own loads or LDS reads in flight (a caller that reuses the LDS
// meets the other waves at a barrier first).  When K
// is not a multiple of 64 the first stage holds the K % 64 leading k values
// (staged through registers, zero-padded) and every later stage is whole.
template <int WM, int WN, int TM, int TN, bool TA, bool TB, bool EDGE = false>
__device__ __forceinline__ void bf16_tile_dma(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                              float* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb,
                                              int64_t ldc, float alpha, float beta, int m0, int n0, __bf16* smem) {
    constexpr int NW = WM * WN;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    typedef DmaOperand<BM, !TA, NW> OpA;  // op(A) rows are k-contiguous unless A is transposed
    typedef DmaOperand<BN, TB, NW> OpB;   // op(B)^T rows are k-contiguous when B is transposed
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    char* const sb = (char*)smem;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w / WN, wn = w % WN;

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int kh = K % 64;  // k values of a short first stage (0: none)
    const int KT = K / 64 + (kh != 0);
    // first k of stage s
    auto kbase = [&](int st) { return st == 0 ? 0 : st * 64 - (kh ? 64 - kh : 0); };
    bf16x8 a0[TM], b0[TN], a1[TM], b1[TN];
    auto frags = [&](const char* stage, int kk, bf16x8(&a)[TM], bf16x8(&b)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = OpA::frag(stage, wm * TM * 32 + i * 32, kk, lane);
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = OpB::frag(stage + OpA::BYTES, wn * TN * 32 + j * 32, kk, lane);
    };
    auto mma = [&](const bf16x8(&a)[TM], const bf16x8(&b)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    };
    // one MFMA, then one LDS read and one global load if any remain, ...
    auto interleave = [&]() {
#pragma unroll
        for (int i = 0; i < TM * TN; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load (to LDS)
        }
    };

    constexpr int LOADS = OpA::PER_WAVE + OpB::PER_WAVE;  // loads per wave per stage
    auto issue = [&](int st, char* buf) {
        OpA::template issue<EDGE>(A, lda, m0, kbase(st), buf, w, lane, M);
        OpB::template issue<EDGE>(B, ldb, n0, kbase(st), buf + OpA::BYTES, w, lane, N);
    };
    if (kh) {
        OpA::template stage_head<EDGE>(A, lda, m0, 0, kh, sb, tid, NW * 64, M);
        OpB::template stage_head<EDGE>(B, ldb, n0, 0, kh, sb + OpA::BYTES, tid, NW * 64, N);
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
    frags(sb, 0, a0, b0);

    // Software pipeline over the four 16-deep steps of a stage: the
    // fragments of step s+1 are read while the MFMAs of step s run.  Before
    // the last step each wave waits for the next stage and for its own reads
    // of this one, and the waves meet at a barrier; this stage's buffer is
    // then free, and the loads of the stage after next go into it while the
    // last step's MFMAs run under the reads of the next stage's first
    // fragments.  A stage's loads thus have a whole iteration of matrix work
    // to land.  Past the last stage the loads repeat it (unused), so every
    // iteration is the same scheduling regions.
    for (int t = 0; t < KT; ++t) {
        char* cur = sb + (t & 1) * STAGE;
        const char* nxt = sb + ((t + 1) & 1) * STAGE;
        __builtin_amdgcn_sched_barrier(0);
        frags(cur, 16, a1, b1);
        mma(a0, b0);
        interleave();
        frags(cur, 32, a0, b0);
        mma(a1, b1);
        interleave();
        frags(cur, 48, a1, b1);
        mma(a0, b0);
        interleave();
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the next stage in; this wave's reads of this one done
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t + 2 < KT ? t + 2 : KT - 1, cur);
        frags(nxt, 0, a0, b0);
        mma(a1, b1);
        interleave();
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
}

// Variant with 32-deep stages (64-byte rows of k-contiguous operands) in R
// LDS slots, for tiles whose 64-deep double buffer would not leave room for
// two work-groups per CU: a 256 x 128 tile takes 48 KiB with R = 2.
template <int WM, int WN, int TM, int TN, int R>
struct Bf16RingShape {
    static constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static constexpr size_t SMEM = (size_t)R * (BM + BN) * 32 * sizeof(__bf16);
};

// The BM x BN tile at (m0, n0) of C = alpha*op(A)op(B) + beta*C, for a tile
// wholly inside M x N, K >= 32 a multiple of 16, 16-byte aligned rows; all
// WM*WN*64 threads of the group call it and each wave returns with none of
// its loads or LDS reads in flight.  When K is not a multiple of 32 the
// first stage holds the K % 32 = 16 leading k values (staged through
// registers, zero-padded).  Its products are bf16_tile's and
// bf16_tile_dma's, 16-deep steps in k order ([0, 16), [16, 32), ...), so the
// three agree bit for bit.
template <int WM, int WN, int TM, int TN, int R, bool TA, bool TB>
__device__ __forceinline__ void bf16_tile_dma_ring(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                                   float* __restrict__ C, int M, int N, int K, int64_t lda,
                                                   int64_t ldb, int64_t ldc, float alpha, float beta, int m0, int n0,
                                                   __bf16* smem) {
    static_assert(R >= 2, "at least two slots");
    constexpr int NW = WM * WN;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    typedef DmaOperand<BM, !TA, NW, __bf16, 64> OpA;
    typedef DmaOperand<BN, TB, NW, __bf16, 64> OpB;
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    char* const sb = (char*)smem;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w / WN, wn = w % WN;

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int kh = K % 32;  // k values of a short first stage (0: none)
    const int KT = K / 32 + (kh != 0);
    auto kbase = [&](int st) { return st == 0 ? 0 : st * 32 - (kh ? 32 - kh : 0); };
    auto slot = [&](int st) { return sb + (st % R) * STAGE; };
    bf16x8 a0[TM], b0[TN], a1[TM], b1[TN];
    auto frags = [&](const char* stage, int kk, bf16x8(&a)[TM], bf16x8(&b)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = OpA::frag(stage, wm * TM * 32 + i * 32, kk, lane);
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = OpB::frag(stage + OpA::BYTES, wn * TN * 32 + j * 32, kk, lane);
    };
    auto mma = [&](const bf16x8(&a)[TM], const bf16x8(&b)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    };
    auto interleave = [&]() {
#pragma unroll
        for (int i = 0; i < TM * TN; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load (to LDS)
        }
    };

    constexpr int LOADS = OpA::PER_WAVE + OpB::PER_WAVE;  // loads per wave per stage
    // stage st into its slot (past the last, the last again: every
    // iteration issues the same loads and waits for the same counts)
    auto issue = [&](int st) {
        const int s = st < KT ? st : KT - 1;
        OpA::issue(A, lda, m0, kbase(s), slot(st), w, lane);
        OpB::issue(B, ldb, n0, kbase(s), slot(st) + OpA::BYTES, w, lane);
    };
    if (kh) {
        OpA::stage_head(A, lda, m0, 0, kh, sb, tid, NW * 64);
        OpB::stage_head(B, ldb, n0, 0, kh, sb + OpA::BYTES, tid, NW * 64);
    } else {
        issue(0);
    }
#pragma unroll
    for (int st = 1; st < R; ++st) issue(st);
    __builtin_amdgcn_s_waitcnt(waitcnt_imm(LOADS * (R - 1), 0));  // stage 0 in LDS
    __builtin_amdgcn_s_barrier();
    frags(slot(0), 0, a0, b0);

    // Iteration t: the second step's fragments are read under the first
    // step's MFMAs; each wave then waits for stage t + 1 (later stages may
    // still be in flight) and for its own reads of stage t, and the waves
    // meet at a barrier.  Stage t's slot is then free, and the loads of
    // stage t + R go into it under the second step's MFMAs and the reads of
    // stage t + 1's first fragments.
    for (int t = 0; t < KT; ++t) {
        __builtin_amdgcn_sched_barrier(0);
        frags(slot(t), 16, a1, b1);
        mma(a0, b0);
        interleave();
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(LOADS * (R - 2), 0));
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t + R);
        frags(slot(t + 1), 0, a0, b0);
        mma(a1, b1);
        interleave();
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
                float* p = C + (int64_t)row * ldc + col;
                float v = alpha * acc[i][j][r];
                if (beta != 0.0f) v += beta * gload<float>(p);
                gstore(p, v);
            }
}

}  // namespace gat
