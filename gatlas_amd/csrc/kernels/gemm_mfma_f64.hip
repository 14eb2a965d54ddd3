// FP64 GEMM on gfx950 matrix cores (v_mfma_f64_16x16x4_f64), kept FP64 end
// to end.  Same work decomposition as the FP32 MFMA kernel: WM x WN waves,
// each owning TM x TN 16x16 output blocks; K staged through double-buffered
// LDS with both operands stored k-contiguous ([m][k] and [n][k], row stride
// BK+2 doubles) so that one ds_read_b128 hands a lane the operands of two
// MFMAs.  MFMA s of a k-group takes, in lane l, k = kk + 2*(l/16) + s.
// Output block layout (measured on gfx950; differs from the FP32 16x16x4
// form): lane l, accumulator r holds row 4*r + l/16 of column l%16.
// WK > 1 splits each staged K tile between WK groups of WM x WN waves, as
// the FP32 k-split tiles do: group g takes the g-th run of the stage's
// 8-deep steps, the groups exchange partial sums through LDS after the K
// loop, and group g adds the WK partials of accumulator register g*4/WK..
// in group order and stores those outputs.
#include "../kernels.hpp"
#include "mfma_common.hpp"

namespace gat {

typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void tile_coords_f64(int wg, int nwg, int tiles_m, int tiles_n, int group_m,
                                                int& tm, int& tn) {
    if ((nwg & 7) == 0) wg = (wg & 7) * (nwg >> 3) + (wg >> 3);
    const int per_group = group_m * tiles_n;
    const int g = wg / per_group;
    const int first_m = g * group_m;
    int gm = tiles_m - first_m;
    if (gm > group_m) gm = group_m;
    const int r = wg - g * per_group;
    tm = first_m + r % gm;
    tn = r / gm;
}

// ROWS x BK tile of doubles: direct (k-contiguous source) or transposed
// (row-contiguous source), stored to LDS as [row][k] with stride BK+2.
template <int ROWS, int BK, int NT, bool TRANS>
struct TileIO64 {
    static constexpr int NV = ROWS * BK / 2 / NT;  // double2 per thread
    static_assert(NV >= 1 && NV * 2 * NT == ROWS * BK, "tile must split evenly into double2 per thread");
    static constexpr int LDK = BK + 2;

    __device__ static __forceinline__ void load(d2v (&reg)[NV], const double* __restrict__ src, int64_t ld,
                                                int r0, int k0, int rows_lim, int k_lim, bool full,
                                                int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int idx = q * NT + tid;
            if (!TRANS) {
                const int r = idx / (BK / 2), kq = (idx % (BK / 2)) * 2;
                const int gr = r0 + r, gk = k0 + kq;
                if (full) {
                    reg[q] = *(const d2v*)(src + (int64_t)gr * ld + gk);
                } else {
#pragma unroll
                    for (int e = 0; e < 2; ++e)
                        reg[q][e] = (gr < rows_lim && gk + e < k_lim) ? src[(int64_t)gr * ld + gk + e] : 0.0;
                }
            } else {
                const int k = idx % BK, rq = (idx / BK) * 2;
                const int gk = k0 + k, gr = r0 + rq;
                if (full) {
                    reg[q] = *(const d2v*)(src + (int64_t)gk * ld + gr);
                } else {
#pragma unroll
                    for (int e = 0; e < 2; ++e)
                        reg[q][e] = (gk < k_lim && gr + e < rows_lim) ? src[(int64_t)gk * ld + gr + e] : 0.0;
                }
            }
        }
    }

    __device__ static __forceinline__ void store(const d2v (&reg)[NV], double* lds, int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int idx = q * NT + tid;
            if (!TRANS) {
                const int r = idx / (BK / 2), kq = (idx % (BK / 2)) * 2;
                *(d2v*)(lds + r * LDK + kq) = reg[q];
            } else {
                const int k = idx % BK, rq = (idx / BK) * 2;
                lds[(rq + 0) * LDK + k] = reg[q][0];
                lds[(rq + 1) * LDK + k] = reg[q][1];
            }
        }
    }
};

// Epilogue of an FP64 macro tile: C = alpha*acc + beta*C for the wave's
// TM x TN 16x16 blocks.  With WK > 1 group g owns accumulator registers
// [g*RS, (g+1)*RS) of each block and adds the WK partial sums of them, sent
// through LDS (the stage buffers, free after the caller's final barrier),
// in group order.  CHECKED false: the tile lies wholly inside M x N.
template <int WM, int WN, int TM, int TN, int WK, bool CHECKED>
__device__ __forceinline__ void dgemm_finish(d4v (&acc)[TM][TN], double* __restrict__ C, int M, int N, int64_t ldc,
                                             double alpha, double beta, int m0, int n0, int wm, int wn, int wmn,
                                             int wk, int lane, double* smemd) {
    constexpr int RS = 4 / WK;  // accumulator registers each group finishes
    static_assert(RS * WK == 4, "each wave group must own whole accumulator registers");
    if constexpr (WK > 1) {
        // partial sums of group g for owner o != g (the stage buffers are free)
        auto slot = [&](int g, int o, int i, int j, int r) {
            const int region = g * (WK - 1) + (o < g ? o : o - 1);
            return ((((region * WM * WN + wmn) * TM + i) * TN + j) * RS + r % RS) * 64 + lane;
        };
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (r / RS != wk) smemd[slot(wk, r / RS, i, j, r)] = acc[i][j][r];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (r / RS == wk) {
                        double v = wk == 0 ? acc[i][j][r] : smemd[slot(0, wk, i, j, r)];
#pragma unroll
                        for (int g = 1; g < WK; ++g) v += g == wk ? acc[i][j][r] : smemd[slot(g, wk, i, j, r)];
                        acc[i][j][r] = v;
                    }
    }

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (WK > 1 && r / RS != wk) continue;
                const int row = m0 + wm * TM * 16 + i * 16 + r * 4 + (lane >> 4);
                const int col = n0 + wn * TN * 16 + j * 16 + (lane & 15);
                if (!CHECKED || (row < M && col < N)) {
                    double* p = C + (int64_t)row * ldc + col;
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v += beta * *p;
                    *p = v;
                }
            }
}

template <int WM, int WN, int TM, int TN, int BK, int WK>
struct DgemmShape {
    static constexpr int BM = WM * TM * 16, BN = WN * TN * 16;
    static constexpr size_t STAGES = 2 * (size_t)(BM + BN) * (BK + 2) * sizeof(double);
    // partial sums the WK wave groups exchange after the K loop
    static constexpr size_t RED = (size_t)(WK - 1) * WM * WN * TM * TN * 256 * sizeof(double);
    static constexpr size_t SMEM = STAGES > RED ? STAGES : RED;
};

// ---- direct-to-LDS staging (BK = 16, interior tiles) ----
// One 16-deep stage is 128 bytes of k per operand row.  A global_load_lds of
// a wavefront writes 1 KiB contiguous in LDS, lane l's 16 bytes at
// base + 16*l.
//   k-contiguous operand: [row][k], chunk c (two doubles) of row r at
//     c ^ ((r >> 1) & 7): the 16 rows a 16-lane group reads at one chunk
//     hit 16 different 16-byte bank slots.  A fragment is one 16-byte read.
//   row-contiguous operand: [k][row] as in memory; chunk c of k-row k at
//     c ^ (8 * ((k >> 1) & 1)).  A fragment is two 8-byte reads, k-rows
//     kk + 2*(lane/16) + s; the shift puts the k-rows of lane groups 0/2 and
//     1/3 in different halves of a 256-byte bank pass.
typedef __attribute__((address_space(3))) void lds_void64_t;
typedef __attribute__((address_space(1))) void glb_void64_t;

template <int ROWS, int NW>
struct DmaColsF64 {
    static constexpr int BYTES = ROWS * 128;
    static constexpr int BLOCKS = BYTES / 1024;  // one wavefront instruction each (8 rows)
    static constexpr int PER_WAVE = BLOCKS / NW;
    static_assert(PER_WAVE >= 1 && PER_WAVE * NW == BLOCKS, "blocks must split evenly over the wavefronts");
    __device__ static __forceinline__ int kswz(int r) { return (r >> 1) & 7; }
    __device__ static __forceinline__ void issue(const double* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int b = i * NW + w;
            const int row = 8 * b + (lane >> 3);
            const double* g = src + (int64_t)(r0 + row) * ld + k0 + 2 * ((lane & 7) ^ kswz(row));
            __builtin_amdgcn_global_load_lds((glb_void64_t*)g, (lds_void64_t*)(lds + b * 1024), 16, 0, 0);
        }
    }
    // operands of the two MFMAs of the 8-deep step at kk, row rbase + lane%16:
    // k = kk + 2*(lane/16) + s
    __device__ static __forceinline__ d2v frag(const char* lds, int rbase, int kk, int lane) {
        const int row = rbase + (lane & 15);
        const int c = (kk >> 1) + (lane >> 4);
        return *(const d2v*)(lds + row * 128 + 16 * (c ^ kswz(row)));
    }
};

template <int ROWS, int NW>
struct DmaRowsF64 {
    static constexpr int CPR = ROWS / 2;         // 16-byte chunks per k-row
    static constexpr int BYTES = 16 * ROWS * 8;  // one 16-deep stage
    static constexpr int BLOCKS = BYTES / 1024;
    static constexpr int PER_WAVE = BLOCKS / NW;
    static constexpr int KPB = 64 / CPR;  // k-rows per block
    static_assert(PER_WAVE >= 1 && PER_WAVE * NW == BLOCKS, "blocks must split evenly over the wavefronts");
    static_assert(CPR >= 16 && CPR <= 64 && 64 % CPR == 0, "row-contiguous FP64 tile: 32..128 rows");
    __device__ static __forceinline__ int rsw(int k) { return 8 * ((k >> 1) & 1); }
    __device__ static __forceinline__ void issue(const double* src, int64_t ld, int r0, int k0, char* lds, int w,
                                                 int lane) {
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int b = i * NW + w;
            const int k = b * KPB + lane / CPR;
            const double* g = src + (int64_t)(k0 + k) * ld + r0 + 2 * ((lane % CPR) ^ rsw(k));
            __builtin_amdgcn_global_load_lds((glb_void64_t*)g, (lds_void64_t*)(lds + b * 1024), 16, 0, 0);
        }
    }
    __device__ static __forceinline__ d2v frag(const char* lds, int rbase, int kk, int lane) {
        const int n = rbase + (lane & 15);
        const int k = kk + 2 * (lane >> 4);  // even: k and k + 1 share the shift
        const double* p = (const double*)(lds + k * (ROWS * 8) + 16 * ((n >> 1) ^ rsw(k))) + (n & 1);
        return d2v{p[0], p[ROWS]};
    }
};

template <int ROWS, bool KCONTIG, int NW>
struct DmaF64Operand;
template <int ROWS, int NW>
struct DmaF64Operand<ROWS, true, NW> : DmaColsF64<ROWS, NW> {};
template <int ROWS, int NW>
struct DmaF64Operand<ROWS, false, NW> : DmaRowsF64<ROWS, NW> {};

// Configurations whose interior tiles take the direct-to-LDS routine
template <int WM, int WN, int TM, int TN, int BK, int WK>
constexpr bool dgemm_has_dma() {
    constexpr int NW = WM * WN * WK, BM = WM * TM * 16, BN = WN * TN * 16;
    return BK == 16 && (WK == 1 || WK == 2) && BM >= 32 && BN >= 32 && BM <= 128 && BN <= 128 &&
           (BM / 8) % NW == 0 && (BN / 8) % NW == 0 && (BM / 8) >= NW && (BN / 8) >= NW;
}

// The tile at (m0, n0), wholly inside M x N, K a positive multiple of 16,
// 16-byte aligned operand rows, staged straight into LDS; the MFMA order is
// the register path's (k = kk + 2*(lane/16) + s per step, steps in k order,
// group g of a k-split tile taking step g of each stage), so the bits agree.
template <int WM, int WN, int TM, int TN, int WK, bool TA, bool TB>
__device__ __forceinline__ void dgemm_tile_dma(const double* __restrict__ A, const double* __restrict__ B,
                                               double* __restrict__ C, int M, int N, int K, int64_t lda,
                                               int64_t ldb, int64_t ldc, double alpha, double beta, int m0, int n0,
                                               double* smemd) {
    constexpr int NW = WM * WN * WK;
    constexpr int BM = WM * TM * 16, BN = WN * TN * 16;
    constexpr int NS = 2 / WK;  // 8-deep steps per wave and stage
    typedef DmaF64Operand<BM, !TA, NW> OpA;
    typedef DmaF64Operand<BN, TB, NW> OpB;
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    char* const sb = (char*)smemd;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wmn = WK == 1 ? w : w % (WM * WN), wk = WK == 1 ? 0 : w / (WM * WN);
    const int wm = wmn / WN, wn = wmn % WN;
    const int kb = WK == 1 ? 0 : wk * NS * 8;

    d4v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;

    const int KT = K / 16;
    d2v a0[TM], b0[TN], a1[TM], b1[TN];
    auto frags = [&](const char* stage, int kk, d2v(&a)[TM], d2v(&b)[TN]) {
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = OpA::frag(stage, wm * TM * 16 + i * 16, kk, lane);
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = OpB::frag(stage + OpA::BYTES, wn * TN * 16 + j * 16, kk, lane);
    };
    auto mma = [&](const d2v(&a)[TM], const d2v(&b)[TN]) {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
    };
    auto interleave = [&]() {
#pragma unroll
        for (int i = 0; i < TM * TN * 2; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load (to LDS)
        }
    };

    constexpr int LOADS = OpA::PER_WAVE + OpB::PER_WAVE;  // loads per wave per stage
    auto issue = [&](int t, char* buf) {
        OpA::issue(A, lda, m0, 16 * t, buf, w, lane);
        OpB::issue(B, ldb, n0, 16 * t, buf + OpA::BYTES, w, lane);
    };
    issue(0, sb);
    if (KT > 1) {
        issue(1, sb + STAGE);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(LOADS, 0));  // stage 0 in LDS
    } else {
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));
    }
    __builtin_amdgcn_s_barrier();
    frags(sb, kb, a0, b0);
    // Before a stage's last step each wave waits for the next stage and its
    // own reads of this one, and the waves meet at a barrier; this stage's
    // buffer is then free, and the loads of the stage after next go into it
    // under the last step's MFMAs (a whole iteration to land).  Past the
    // last stage the loads repeat it (unused).
    for (int t = 0; t < KT; ++t) {
        char* cur = sb + (t & 1) * STAGE;
        const char* nxt = sb + ((t + 1) & 1) * STAGE;
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (NS == 2) {
            frags(cur, kb + 8, a1, b1);
            mma(a0, b0);
            interleave();
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the next stage in; this wave's reads of this one done
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
            issue(t + 2 < KT ? t + 2 : KT - 1, cur);
            frags(nxt, kb, a0, b0);
            mma(a1, b1);
            interleave();
        } else {
            mma(a0, b0);
            interleave();
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
            issue(t + 2 < KT ? t + 2 : KT - 1, cur);
            frags(nxt, kb, a0, b0);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the repeated loads land before LDS is reused
    if constexpr (WK > 1) __syncthreads();  // the exchange reuses the stage buffers
    dgemm_finish<WM, WN, TM, TN, WK, false>(acc, C, M, N, ldc, alpha, beta, m0, n0, wm, wn, wmn, wk, lane, smemd);
}

template <int WM, int WN, int TM, int TN, int BK, int WK, bool TA, bool TB>
__global__ __launch_bounds__(WM* WN * WK * 64) void dgemm_mfma_kernel(
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C, int M, int N, int K,
    int64_t lda, int64_t ldb, int64_t ldc, int64_t sA, int64_t sB, int64_t sC, double alpha, double beta,
    int tiles_m, int tiles_n, int group_m, int vec_ok) {
    constexpr int NT = WM * WN * WK * 64;
    constexpr int BM = WM * TM * 16, BN = WN * TN * 16;
    constexpr int LDK = BK + 2;
    static_assert(BK % 8 == 0, "BK must cover whole k-groups of 8");
    constexpr int NS = BK / 8 / WK;  // 8-deep steps per wave and stage
    static_assert(NS * 8 * WK == BK, "the k-split groups must share a stage's steps evenly");
    constexpr int RS = 4 / WK;  // accumulator registers each group finishes
    static_assert(RS * WK == 4, "each wave group must own whole accumulator registers");
    typedef TileIO64<BM, BK, NT, TA> IoA;
    typedef TileIO64<BN, BK, NT, !TB> IoB;

    extern __shared__ double smemd[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wmn = wave % (WM * WN), wk = wave / (WM * WN);  // wk = 0 when WK == 1
    const int wm = wmn / WN, wn = wmn % WN;

    A += (int64_t)blockIdx.y * sA;
    B += (int64_t)blockIdx.y * sB;
    C += (int64_t)blockIdx.y * sC;

    int tmi, tni;
    tile_coords_f64(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    const int m0 = tmi * BM, n0 = tni * BN;
    const bool full_mn = vec_ok && (m0 + BM <= M) && (n0 + BN <= N);
    if constexpr (dgemm_has_dma<WM, WN, TM, TN, BK, WK>()) {
        if (full_mn && K >= 16 && K % 16 == 0) {
            dgemm_tile_dma<WM, WN, TM, TN, WK, TA, TB>(A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, m0, n0, smemd);
            return;
        }
    }

    d4v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;

    d2v ra[IoA::NV], rb[IoB::NV];
    const int KT = (K + BK - 1) / BK;

    auto gload = [&](int t) {
        const int k0 = t * BK;
        const bool full = full_mn && (k0 + BK <= K);
        IoA::load(ra, A, lda, m0, k0, M, K, full, tid);
        IoB::load(rb, B, ldb, n0, k0, N, K, full, tid);
    };

    gload(0);
    IoA::store(ra, smemd, tid);
    IoB::store(rb, smemd + BM * LDK, tid);
    __syncthreads();

    const int arow = wm * TM * 16 + (lane & 15);
    const int brow = wn * TN * 16 + (lane & 15);
    const int kq = 2 * (lane >> 4) + wk * NS * 8;  // this k-group's run of steps

    for (int t = 0; t < KT; ++t) {
        const double* As = smemd + (t & 1) * (BM + BN) * LDK;
        const double* Bs = As + BM * LDK;
        if (t + 1 < KT) gload(t + 1);
#pragma unroll
        for (int kk = 0; kk < NS * 8; kk += 8) {
            d2v a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = *(const d2v*)(As + (arow + i * 16) * LDK + kk + kq);
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = *(const d2v*)(Bs + (brow + j * 16) * LDK + kk + kq);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
        }
        if (t + 1 < KT) {
            double* nA = smemd + ((t + 1) & 1) * (BM + BN) * LDK;
            IoA::store(ra, nA, tid);
            IoB::store(rb, nA + BM * LDK, tid);
        }
        __syncthreads();
    }

    dgemm_finish<WM, WN, TM, TN, WK, true>(acc, C, M, N, ldc, alpha, beta, m0, n0, wm, wn, wmn, wk, lane, smemd);
}

// Instantiated configurations: (WM, WN, TM, TN, BK, WK); the WK > 1 entries
// are for problems whose tiles give fewer than two work-groups per CU.
#define GAT_F64_MFMA_CFGS(X) \
    X(2, 2, 2, 2, 16, 1)     \
    X(2, 2, 4, 4, 16, 1)     \
    X(2, 2, 4, 2, 16, 1)     \
    X(2, 2, 2, 2, 32, 1)     \
    X(2, 2, 4, 4, 8, 1)      \
    X(1, 1, 2, 2, 32, 1)     \
    X(2, 2, 2, 2, 16, 2)     \
    X(2, 2, 2, 2, 32, 2)     \
    X(2, 2, 2, 2, 32, 4)     \
    X(4, 4, 2, 2, 16, 1)     \
    X(4, 2, 2, 2, 16, 1)     \
    X(2, 4, 2, 2, 16, 1)

static const MfmaCfg kF64Cfgs[] = {
#define X(wm, wn, tm, tn, bk, wk) {wm, wn, tm, tn, bk, 16, wk},
    GAT_F64_MFMA_CFGS(X)
#undef X
};

int mfma_f64_num_cfgs() { return (int)(sizeof(kF64Cfgs) / sizeof(kF64Cfgs[0])); }
MfmaCfg mfma_f64_cfg(int i) { return kF64Cfgs[i]; }

bool mfma_f64_has(const MfmaCfg& c) {
    for (const MfmaCfg& k : kF64Cfgs)
        if (k.waves_m == c.waves_m && k.waves_n == c.waves_n && k.tm == c.tm && k.tn == c.tn &&
            k.bk == c.bk && k.blk == c.blk && k.wk == c.wk)
            return true;
    return false;
}

template <int WM, int WN, int TM, int TN, int BK, int WK>
static int launch_one64(const GemmArgs& a, hipStream_t st) {
    constexpr int BM = WM * TM * 16, BN = WN * TN * 16;
    constexpr size_t smem_reg = DgemmShape<WM, WN, TM, TN, BK, WK>::SMEM;
    constexpr size_t smem_dma = dgemm_has_dma<WM, WN, TM, TN, BK, WK>() ? 2 * (size_t)(BM + BN) * 128 : 0;
    constexpr size_t smem = smem_reg > smem_dma ? smem_reg : smem_dma;
    const int tiles_m = (int)((a.M + BM - 1) / BM), tiles_n = (int)((a.N + BN - 1) / BN);
    const double* A = (const double*)a.A;
    const double* B = (const double*)a.B;
    const int vec_ok = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && (a.lda % 2 == 0) &&
                       (a.ldb % 2 == 0) && (a.strideA % 2 == 0) && (a.strideB % 2 == 0);
    typedef void (*kfn)(const double*, const double*, double*, int, int, int, int64_t, int64_t, int64_t,
                        int64_t, int64_t, int64_t, double, double, int, int, int, int);
    kfn f;
    if (!a.transA && !a.transB) f = dgemm_mfma_kernel<WM, WN, TM, TN, BK, WK, false, false>;
    else if (a.transA && !a.transB) f = dgemm_mfma_kernel<WM, WN, TM, TN, BK, WK, true, false>;
    else if (!a.transA && a.transB) f = dgemm_mfma_kernel<WM, WN, TM, TN, BK, WK, false, true>;
    else f = dgemm_mfma_kernel<WM, WN, TM, TN, BK, WK, true, true>;
    if (smem > 65536) GAT_HIP_TRY(allow_dynamic_lds((const void*)f, smem));
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)a.batch);
    hipLaunchKernelGGL(f, grid, dim3(WM * WN * WK * 64), smem, st, A, B, (double*)a.C, (int)a.M, (int)a.N,
                       (int)a.K, a.lda, a.ldb, a.ldc, a.strideA, a.strideB, a.strideC, a.alpha, a.beta,
                       tiles_m, tiles_n, group_m, vec_ok);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_mfma_f64(const MfmaCfg& c, const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) return ST_INVALID;
    if (c.blk != 16) {
        set_error("mfma f64: only 16x16 blocks exist");
        return ST_UNSUPPORTED;
    }
#define X(p_wm, p_wn, p_tm, p_tn, p_bk, p_wk)                                                        \
    if (c.waves_m == p_wm && c.waves_n == p_wn && c.tm == p_tm && c.tn == p_tn && c.bk == p_bk &&    \
        c.wk == p_wk)                                                                                 \
        return launch_one64<p_wm, p_wn, p_tm, p_tn, p_bk, p_wk>(a, st);
    GAT_F64_MFMA_CFGS(X)
#undef X
    set_error("mfma f64: configuration not instantiated");
    return ST_UNSUPPORTED;
}

}  // namespace gat
