// Direct-to-LDS variant of the FP32 matrix-core macro tile (BK = 32) for
// interior tiles: both operands go from HBM into LDS with gfx950's 16-byte
// global_load_lds, so no register holds a staged tile, no LDS store
// instruction runs, and the next K stage lands in LDS while the MFMAs run
// on the current one.  The products are the register-staged tile's
// (sgemm_tile.hpp, 32x32x2 blocks): lane l takes k = kk + 4*(l/32) + s in
// MFMA s of an 8-deep step, in the same order, so the two paths agree bit
// for bit.
//
// Layouts of one 32-deep stage in LDS:
//   k-contiguous operand (op(A) rows when A is not transposed, op(B)^T rows
//     when B is transposed): [row][k], 128 bytes (32 floats) per row,
//     16-byte chunk c of row r at c ^ ((r >> 1) & 7); a fragment (four k
//     values of one row) is one 16-byte read.  Rows r and r + 1 fall in the
//     two halves of a 256-byte pass over the banks and rows r, r + 2, ...
//     differ in the swizzle, so the 16 rows one read covers per 16 lanes
//     hit 16 different 16-byte bank slots (the c ^ (r & 7) swizzle of the
//     BF16 layout leaves rows r and r + 8 on one slot).  B's rows keep
//     c ^ (r & 7) when A is transposed, which measured faster there.
//   row-contiguous operand (A transposed, B not transposed): [k][row] as in
//     memory, ROWS floats per k-row; 16-byte chunk c of k-row k sits at
//     c ^ (8 * ((k >> 2) & 1)).  A fragment is four 4-byte reads, k-rows
//     kk + s (lanes 0..31) and kk + 4 + s (lanes 32..63); k-rows are a
//     multiple of 256 bytes long, so without the swizzle the two halves of
//     the wavefront would read the same 32 banks.
#pragma once

#include "bf16_dma_tile.hpp"
#include "sgemm_tile.hpp"

namespace gat {

// SB selects the k bit that moves a k-row's chunks by 128 bytes: bit 2
// for the FP32 fragments (k-rows kk + s and kk + 4 + s, read together),
// bit 3 for the split-BF16x3 ones (k-rows kk + s and kk + 8 + s).
template <int ROWS, int NW, int SB = 2>
struct DmaRowsF32 {
    static constexpr int CPR = ROWS / 4;            // 16-byte chunks per k-row
    static constexpr int BYTES = 32 * ROWS * 4;     // one 32-deep stage
    static constexpr int BLOCKS = BYTES / 1024;     // one wavefront instruction each
    static constexpr int PER_WAVE = BLOCKS / NW;
    static constexpr int KPB = 64 / CPR;            // k-rows per block
    static_assert(PER_WAVE * NW == BLOCKS, "blocks must split evenly over the wavefronts");
    static_assert(CPR >= 16 && CPR <= 64 && 64 % CPR == 0, "row-contiguous FP32 tile: 64..256 rows");

    __device__ static __forceinline__ int rsw(int k) { return 8 * ((k >> SB) & 1); }

    // First row (column of the stored [k][row] matrix) of the 16-byte chunk
    // at r; a tile that reaches past the last row (EDGE, rlim rows in all, a
    // multiple of 4) fetches the last whole chunk instead, so every load
    // stays inside the matrix and the rows past it hold copies.
    template <bool EDGE>
    __device__ static __forceinline__ int chunk_row(int r, int rlim) {
        return EDGE ? (r < rlim - 4 ? r : rlim - 4) : r;
    }

    // Wavefront w fetches its blocks of the stage starting at k0 into `lds`.
    template <bool EDGE = false>
    __device__ static __forceinline__ void issue(const float* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane, int rlim = 0) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int b = i * NW + w;
            const int k = b * KPB + lane / CPR;
            const float* g = src + (int64_t)(k0 + k) * ld + chunk_row<EDGE>(r0 + 4 * ((lane % CPR) ^ rsw(k)), rlim);
            __builtin_amdgcn_global_load_lds((glb_void_t*)g, (lds_void_t*)(lds + b * 1024), 16, 0, 0);
        }
    }

    // The same loads through per-lane addresses computed once per tile:
    // prep() gives the stage at k0 = 0, and the stage at k0 is then
    // issue_at(p, k0 * ld), one 64-bit add per load.
    template <bool EDGE = false>
    __device__ static __forceinline__ void prep(const float* src, int64_t ld, int r0, int w, int lane,
                                                const float* (&p)[PER_WAVE], int rlim = 0) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int b = i * NW + w;
            const int k = b * KPB + lane / CPR;
            p[i] = src + (int64_t)k * ld + chunk_row<EDGE>(r0 + 4 * ((lane % CPR) ^ rsw(k)), rlim);
        }
    }
    __device__ static __forceinline__ void issue_at(const float* const (&p)[PER_WAVE], int64_t off, char* lds,
                                                    int w) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i)
            __builtin_amdgcn_global_load_lds((glb_void_t*)(p[i] + off), (lds_void_t*)(lds + (i * NW + w) * 1024), 16,
                                             0, 0);
    }
    __device__ static __forceinline__ int64_t stage_off(int k0, int64_t ld) { return (int64_t)k0 * ld; }

    // The k values [k0, k0 + klen) (klen < 32 a multiple of 4) written into
    // the last klen k-rows of a stage in `lds`, the first 32 - klen zero,
    // through registers by all nt work-items: a short first stage.
    template <bool EDGE = false>
    __device__ static __forceinline__ void stage_head(const float* src, int64_t ld, int r0, int k0, int klen,
                                                      char* lds, int tid, int nt, int rlim = 0) {
        for (int q = tid; q < 32 * CPR; q += nt) {
            const int k = q / CPR, c = q % CPR;
            const int kr = k - (32 - klen);  // k value held by k-row k, < 0: padding
            f4v v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (kr >= 0) v = gload<f4v>(src + (int64_t)(k0 + kr) * ld + chunk_row<EDGE>(r0 + 4 * c, rlim));
            *(f4v*)(lds + k * (ROWS * 4) + 16 * (c ^ rsw(k))) = v;
        }
    }

    // B-side operand of MFMA s of the 8-deep step at kk for row rbase + lane%32:
    // element (k = kk + 4*(lane/32) + s, row)
    __device__ static __forceinline__ void frag(const char* lds, int rbase, int kk, int lane, float (&v)[4]) {
        const int n = rbase + (lane & 31);
        const int k = kk + 4 * (lane >> 5);  // kk is a multiple of 8: k .. k+3 share the swizzle
        const float* p = (const float*)(lds + k * (ROWS * 4) + 16 * ((n >> 2) ^ rsw(k))) + (n & 3);
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = p[s * ROWS];
    }

    // BF16 MFMA operand in FP32 (split-BF16x3 tiles, SB = 3): element
    // (k = kk + 8*(lane/32) + e, row rbase + lane%32) in v[e], kk a multiple
    // of 16, so all eight k-rows of a lane share the swizzle
    __device__ static __forceinline__ void frag8(const char* lds, int rbase, int kk, int lane, float (&v)[8]) {
        static_assert(SB == 3, "eight k-rows per lane need the bit-3 swizzle");
        const int n = rbase + (lane & 31);
        const int k = kk + 8 * (lane >> 5);
        const float* p = (const float*)(lds + k * (ROWS * 4) + 16 * ((n >> 2) ^ rsw(k))) + (n & 3);
#pragma unroll
        for (int s = 0; s < 8; ++s) v[s] = p[s * ROWS];
    }
};

// k-contiguous FP32 operand of a 32-deep K stage: ROWS rows of 128 bytes.
template <int ROWS, int NW, int SW>
struct DmaColsF32 {
    static constexpr int BYTES = ROWS * 128;
    static constexpr int BLOCKS = BYTES / 1024;  // one wavefront instruction each (8 rows)
    static constexpr int PER_WAVE = BLOCKS / NW;
    static_assert(PER_WAVE * NW == BLOCKS, "blocks must split evenly over the wavefronts");

    __device__ static __forceinline__ int kswz(int r) { return SW ? (r >> 1) & 7 : r & 7; }

    // Stored row of tile row r0 + row; a tile that reaches past the last
    // row (EDGE, rlim rows in all) reads the last row there instead, so
    // every load stays inside the matrix.
    template <bool EDGE>
    __device__ static __forceinline__ int srow(int r, int rlim) {
        return EDGE ? (r < rlim ? r : rlim - 1) : r;
    }

    template <bool EDGE = false>
    __device__ static __forceinline__ void issue(const float* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane, int rlim = 0) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int b = i * NW + w;
            const int row = 8 * b + (lane >> 3);
            const float* g = src + (int64_t)srow<EDGE>(r0 + row, rlim) * ld + k0 + 4 * ((lane & 7) ^ kswz(row));
            __builtin_amdgcn_global_load_lds((glb_void_t*)g, (lds_void_t*)(lds + b * 1024), 16, 0, 0);
        }
    }

    // per-lane addresses computed once per tile (see DmaRowsF32::prep); the
    // stage at k0 is issue_at(p, k0)
    template <bool EDGE = false>
    __device__ static __forceinline__ void prep(const float* src, int64_t ld, int r0, int w, int lane,
                                                const float* (&p)[PER_WAVE], int rlim = 0) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int row = 8 * (i * NW + w) + (lane >> 3);
            p[i] = src + (int64_t)srow<EDGE>(r0 + row, rlim) * ld + 4 * ((lane & 7) ^ kswz(row));
        }
    }
    __device__ static __forceinline__ void issue_at(const float* const (&p)[PER_WAVE], int64_t off, char* lds,
                                                    int w) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i)
            __builtin_amdgcn_global_load_lds((glb_void_t*)(p[i] + off), (lds_void_t*)(lds + (i * NW + w) * 1024), 16,
                                             0, 0);
    }
    __device__ static __forceinline__ int64_t stage_off(int k0, int64_t /*ld*/) { return k0; }

    // short first stage through registers, k values in the last klen
    // positions of each row (see DmaRowsF32::stage_head)
    template <bool EDGE = false>
    __device__ static __forceinline__ void stage_head(const float* src, int64_t ld, int r0, int k0, int klen,
                                                      char* lds, int tid, int nt, int rlim = 0) {
        for (int q = tid; q < ROWS * 8; q += nt) {
            const int row = q >> 3, c = q & 7;
            const int kr = 4 * c - (32 - klen);  // first k value of chunk c, < 0: padding
            f4v v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (kr >= 0) v = gload<f4v>(src + (int64_t)srow<EDGE>(r0 + row, rlim) * ld + k0 + kr);
            *(f4v*)(lds + row * 128 + 16 * (c ^ kswz(row))) = v;
        }
    }

    // operand of MFMA s of the 8-deep step at kk for row rbase + lane%32:
    // element (row, k = kk + 4*(lane/32) + s)
    __device__ static __forceinline__ void frag(const char* lds, int rbase, int kk, int lane, float (&v)[4]) {
        const int row = rbase + (lane & 31);
        const int c = (kk >> 2) + (lane >> 5);
        const f4v f = *(const f4v*)(lds + row * 128 + 16 * (c ^ kswz(row)));
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = f[s];
    }

    // BF16 MFMA operand in FP32 (split-BF16x3 tiles): element (row,
    // k = kk + 8*(lane/32) + e) in v[e], kk a multiple of 16: two 16-byte reads
    __device__ static __forceinline__ void frag8(const char* lds, int rbase, int kk, int lane, float (&v)[8]) {
        const int row = rbase + (lane & 31);
        const int c = (kk >> 2) + 2 * (lane >> 5);
        const f4v f0 = *(const f4v*)(lds + row * 128 + 16 * (c ^ kswz(row)));
        const f4v f1 = *(const f4v*)(lds + row * 128 + 16 * ((c + 1) ^ kswz(row)));
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            v[s] = f0[s];
            v[s + 4] = f1[s];
        }
    }
};

template <int ROWS, bool KCONTIG, int NW, int SW = 1>
struct DmaF32Operand;

template <int ROWS, int NW, int SW>
struct DmaF32Operand<ROWS, true, NW, SW> {
    typedef DmaColsF32<ROWS, NW, SW> Op;
    static constexpr int BYTES = Op::BYTES, PER_WAVE = Op::PER_WAVE;
    template <bool EDGE = false>
    __device__ static __forceinline__ void issue(const float* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane, int rlim = 0) {
        Op::template issue<EDGE>(src, ld, r0, k0, lds, w, lane, rlim);
    }
    __device__ static __forceinline__ void frag(const char* lds, int rbase, int kk, int lane, float (&v)[4]) {
        Op::frag(lds, rbase, kk, lane, v);
    }
    template <bool EDGE = false>
    __device__ static __forceinline__ void stage_head(const float* src, int64_t ld, int r0, int k0, int klen,
                                                      char* lds, int tid, int nt, int rlim = 0) {
        Op::template stage_head<EDGE>(src, ld, r0, k0, klen, lds, tid, nt, rlim);
    }
};

template <int ROWS, int NW, int SW>
struct DmaF32Operand<ROWS, false, NW, SW> {
    typedef DmaRowsF32<ROWS, NW> Op;
    static constexpr int BYTES = Op::BYTES, PER_WAVE = Op::PER_WAVE;
    template <bool EDGE = false>
    __device__ static __forceinline__ void issue(const float* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane, int rlim = 0) {
        Op::template issue<EDGE>(src, ld, r0, k0, lds, w, lane, rlim);
    }
    __device__ static __forceinline__ void frag(const char* lds, int rbase, int kk, int lane, float (&v)[4]) {
        Op::frag(lds, rbase, kk, lane, v);
    }
    template <bool EDGE = false>
    __device__ static __forceinline__ void stage_head(const float* src, int64_t ld, int r0, int k0, int klen,
                                                      char* lds, int tid, int nt, int rlim = 0) {
        Op::template stage_head<EDGE>(src, ld, r0, k0, klen, lds, tid, nt, rlim);
    }
};

template <int WM, int WN, int TM, int TN, int WK = 1>
struct SgemmDmaShape {
    static constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static constexpr size_t STAGES = 2 * (size_t)(BM + BN) * 32 * sizeof(float);
    static constexpr size_t RED = SgemmShape<WM, WN, TM, TN, 32, 32, WK>::RED;  // k-split exchange
    static constexpr size_t SMEM = STAGES > RED ? STAGES : RED;
};

// Computes the BM x BN tile at (m0, n0) of C = alpha*op(A)op(B) + beta*C for
// a tile wholly inside M x N (EDGE: any tile, with M a multiple of 4 when A
// is transposed and N one when B is not -- rows past the matrix are loaded
// as copies of its last rows and not stored), K >= 32 a multiple of 16,
// 16-byte aligned operand rows; all WM*WN*WK*64 threads of the group call
// it; ends after a barrier.  When K is an odd multiple of 16 the first stage holds the 16
// leading k values in its second half (staged through registers, the first
// half zero) and every later stage is whole, so stage s covers k
// [32s - 16, 32s + 16).  The MFMAs pair the same k values as the register
// path, plus all-zero products ahead of them; with two k-split groups the
// groups swap stage halves, so each still sums the register path's k values
// (its groups take [32t, 32t + 16) and [32t + 16, 32t + 32)) in the same
// order, and the two partial sums commute: the bits agree.  WK > 1 splits each stage's four 8-deep steps between WK groups
// of WM x WN waves as the register-staged k-split tile does (group g takes
// steps g*4/WK ..), with the same partial-sum exchange, so the bits agree.
template <int WM, int WN, int TM, int TN, bool TA, bool TB, int WK = 1, bool EDGE = false>
__device__ __forceinline__ void sgemm_tile_dma(const float* __restrict__ A, const float* __restrict__ B,
                                               float* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb,
                                               int64_t ldc, float alpha, float beta, int m0, int n0, float* smem) {
    constexpr int NW = WM * WN * WK;  // wavefronts issuing the stage loads
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr int NS = 4 / WK;  // 8-deep steps per wave and stage
    static_assert(NS == 4 || NS == 2, "k-split tiles split a 32-deep stage into two or four runs");
    // k-contiguous rows swizzled by (r >> 1) & 7, except B's when A is
    // transposed: measured at 8192^3, TT runs 144 TFLOP/s with r & 7 there
    // against 138 with (r >> 1) & 7, while NT gains 141 -> 143 and NN
    // (A's swizzle) 137.5 -> 139.6 from (r >> 1) & 7.
    typedef DmaF32Operand<BM, !TA, NW, 1> OpA;          // op(A) rows are k-contiguous unless A is transposed
    typedef DmaF32Operand<BN, TB, NW, TA ? 0 : 1> OpB;  // op(B)^T rows are k-contiguous when B is transposed
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    char* const sb = (char*)smem;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wmn = WK == 1 ? w : w % (WM * WN), wk = WK == 1 ? 0 : w / (WM * WN);
    const int wm = wmn / WN, wn = wmn % WN;
    const int kh = K % 32;  // k values of a short first stage (0: none)
    // this k-group's run of steps in a stage (halves swapped with a short first stage)
    const int kb = WK == 1 ? 0 : (kh ? (WK - 1 - wk) : wk) * NS * 8;

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int KT = K / 32 + (kh != 0);
    auto kbase = [&](int st) { return st == 0 ? 0 : st * 32 - (kh ? 32 - kh : 0); };  // first k of stage st
    float a0[TM][4], b0[TN][4], a1[TM][4], b1[TN][4];
    auto frags = [&](const char* stage, int kk, float(&a)[TM][4], float(&b)[TN][4]) {
#pragma unroll
        for (int i = 0; i < TM; ++i) OpA::frag(stage, wm * TM * 32 + i * 32, kk, lane, a[i]);
#pragma unroll
        for (int j = 0; j < TN; ++j) OpB::frag(stage + OpA::BYTES, wn * TN * 32 + j * 32, kk, lane, b[j]);
    };
    auto mma = [&](const float(&a)[TM][4], const float(&b)[TN][4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
    };
    // one MFMA, then one LDS read and one global load if any remain, ...
    // (8192^3 NN 139.4 -> 142.9, TN 144.5 -> 146.7 TFLOP/s against the
    // compiler's own order)
    auto interleave = [&]() {
#pragma unroll
        for (int i = 0; i < TM * TN * 4; ++i) {
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
    frags(sb, kb, a0, b0);

    // Software pipeline over the wave's NS 8-deep steps of a stage: the
    // fragments of step s+1 are read while the MFMAs of step s run.  Before
    // the last step each wave waits for the next stage and for its own
    // reads of this one, and the waves meet at a barrier; this stage's
    // buffer is then free, and the loads of the stage after next go into it
    // while the last step's MFMAs run under the reads of the next stage's
    // first fragments: a stage's loads have a whole iteration to land.  Past
    // the last stage the loads repeat it (unused), so every iteration has
    // the same shape.
    for (int t = 0; t < KT; ++t) {
        char* cur = sb + (t & 1) * STAGE;
        const char* nxt = sb + ((t + 1) & 1) * STAGE;
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (NS == 4) {
            frags(cur, kb + 8, a1, b1);
            mma(a0, b0);
            interleave();
            frags(cur, kb + 16, a0, b0);
            mma(a1, b1);
            interleave();
            frags(cur, kb + 24, a1, b1);
            mma(a0, b0);
            interleave();
        } else {
            frags(cur, kb + 8, a1, b1);
            mma(a0, b0);
            interleave();
        }
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the next stage in; this wave's reads of this one done
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t + 2 < KT ? t + 2 : KT - 1, cur);
        frags(nxt, kb, a0, b0);
        mma(a1, b1);
        interleave();
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the repeated loads land before LDS is reused
    // the k-split exchange reuses the stage buffers: every wave's last
    // fragment reads must be done
    if constexpr (WK > 1) __syncthreads();

    sgemm_finish<WM, WN, TM, TN, 32, WK, EDGE>(acc, C, M, N, ldc, alpha, beta, m0, n0, wm, wn, wmn, wk, lane,
                                               smem);
    if constexpr (WK == 1) __syncthreads();  // LDS free for the caller's next tile
}

}  // namespace gat
