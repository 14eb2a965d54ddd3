// One macro tile of the FP32 matrix-core GEMM (v_mfma_f32_32x32x2_f32 /
// 16x16x4_f32), shared by the per-problem kernel and the persistent grouped
// kernel.
//
//   * macro tile BM x BN = (WM*TM*BLK) x (WN*TN*BLK), WM x WN wavefronts,
//     each wave owns TM x TN output blocks of BLK x BLK in accumulators;
//   * K is staged through double-buffered LDS, BK deep.  Both operands are
//     kept k-contiguous in LDS (A as [m][k], B as [n][k], row stride BK+4
//     floats) so one ds_read_b128 gives a lane the operands of four MFMAs;
//     the +4 pad spreads the 32 rows a half-wave reads over distinct banks;
//   * the next K tile is fetched from HBM into registers while the current
//     one is multiplied, then written to the other LDS buffer: one barrier
//     per K step, and no write-after-read hazard on the LDS tiles;
//   * epilogue C = alpha*acc + beta*C (C untouched when beta == 0), with
//     each half-wave storing 32 consecutive floats of a row;
//   * WK > 1 splits every staged K tile between WK groups of WM x WN waves
//     (group g takes the g-th run of its 8- or 16-deep steps): a small tile
//     then keeps WK waves per SIMD busy instead of one.  After the K loop
//     the groups exchange partial sums through LDS; each adds the WK
//     partials of its share of the outputs in group order and stores it.
// MFMA k-lane mapping: MFMA s of a group takes, in lane l, the k index
// kk + 4*(l/32) + s (32x32x2) or kk + 4*(l/16) + s (16x16x4).  A and B use
// the same mapping, so the product is the full sum over the group's k.
#pragma once

#include "mfma_common.hpp"

namespace gat {

template <int BLK>
struct MfmaF32;

template <>
struct MfmaF32<32> {
    typedef f16v acc_t;
    static constexpr int NREG = 16;
    static constexpr int KGROUP = 8;  // k covered by one b128 read per lane (4 MFMAs)
    __device__ static __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int lane_row(int lane) { return lane & 31; }
    __device__ static __forceinline__ int lane_kq(int lane) { return 4 * (lane >> 5); }
    __device__ static __forceinline__ int out_row(int lane, int r) { return mfma32_row(lane, r); }
    __device__ static __forceinline__ int out_col(int lane) { return lane & 31; }
};

template <>
struct MfmaF32<16> {
    typedef f4v acc_t;
    static constexpr int NREG = 4;
    static constexpr int KGROUP = 16;
    __device__ static __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int lane_row(int lane) { return lane & 15; }
    __device__ static __forceinline__ int lane_kq(int lane) { return 4 * (lane >> 4); }
    __device__ static __forceinline__ int out_row(int lane, int r) { return (lane >> 4) * 4 + r; }
    __device__ static __forceinline__ int out_col(int lane) { return lane & 15; }
};

// Loads a ROWS x BK tile of a k-contiguous ("direct") or row-contiguous
// ("transposed") source into registers, and stores it to LDS as [row][k].
//  direct:     element (r, k) at src[(r0 + r) * ld + k0 + k]
//  transposed: element (r, k) at src[(k0 + k) * ld + r0 + r]
// Transposed tiles are walked k-fastest across lanes so that the scalar LDS
// writes of a wavefront hit consecutive banks.
template <int ROWS, int BK, int NT, bool TRANS>
struct TileIO {
    static constexpr int NV = ROWS * BK / 4 / NT;  // float4 per thread
    static_assert(NV >= 1 && NV * 4 * NT == ROWS * BK, "tile must split evenly into float4 per thread");
    static constexpr int LDK = BK + 4;

    // full: the whole tile is in bounds (no checks).  Otherwise each vector
    // is checked on its own: out-of-range elements read as zero, a vector
    // wholly in range is still one 16-byte load when `vec` (aligned source).
    __device__ static __forceinline__ void load(f4v (&reg)[NV], const float* __restrict__ src, int64_t ld,
                                                int r0, int k0, int rows_lim, int k_lim, bool full, bool vec,
                                                int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int idx = q * NT + tid;
            if (!TRANS) {
                const int r = idx / (BK / 4), kq = (idx % (BK / 4)) * 4;
                const int gr = r0 + r, gk = k0 + kq;
                const float* p = src + (int64_t)gr * ld + gk;
                if (full) {  // uniform: interior tiles
                    reg[q] = gload<f4v>(p);
                } else if (vec && gr < rows_lim && gk + 4 <= k_lim) {
                    reg[q] = gload<f4v>(p);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) reg[q][e] = (gr < rows_lim && gk + e < k_lim) ? gload<float>(p + e) : 0.0f;
                }
            } else {
                const int k = idx % BK, rq = (idx / BK) * 4;
                const int gk = k0 + k, gr = r0 + rq;
                const float* p = src + (int64_t)gk * ld + gr;
                if (full) {
                    reg[q] = gload<f4v>(p);
                } else if (vec && gk < k_lim && gr + 4 <= rows_lim) {
                    reg[q] = gload<f4v>(p);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) reg[q][e] = (gk < k_lim && gr + e < rows_lim) ? gload<float>(p + e) : 0.0f;
                }
            }
        }
    }

    __device__ static __forceinline__ void store(const f4v (&reg)[NV], float* lds, int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int idx = q * NT + tid;
            if (!TRANS) {
                const int r = idx / (BK / 4), kq = (idx % (BK / 4)) * 4;
                *(f4v*)(lds + r * LDK + kq) = reg[q];
            } else {
                const int k = idx % BK, rq = (idx / BK) * 4;
#pragma unroll
                for (int e = 0; e < 4; ++e) lds[(rq + e) * LDK + k] = reg[q][e];
            }
        }
    }
};

template <int WM, int WN, int TM, int TN, int BK, int BLK, int WK = 1>
struct SgemmShape {
    static constexpr int NT = WM * WN * WK * 64;
    static constexpr int BM = WM * TM * BLK, BN = WN * TN * BLK;
    static constexpr int LDK = BK + 4;
    static constexpr size_t STAGES = 2 * (size_t)(BM + BN) * LDK * sizeof(float);
    // partial sums the WK wave groups exchange after the K loop: each sends
    // the (WK-1)/WK of its accumulators that other groups finish
    static constexpr size_t RED = (size_t)(WK - 1) * WM * WN * TM * TN * BLK * BLK * sizeof(float);
    static constexpr size_t SMEM = STAGES > RED ? STAGES : RED;
};

// Epilogue of an FP32 macro tile: C = alpha*acc + beta*C for the wave's
// TM x TN blocks (C untouched when beta == 0; outputs outside M x N are
// skipped).  With WK > 1 every wave group finishes NREG/WK of each block's
// accumulator registers: group g owns registers [g*RS, (g+1)*RS) and
// receives the other groups' partial sums of them through LDS (the caller's
// K loop ended with a barrier, so the stage buffers are free), adding the
// WK partials in group order.  The C reads and stores are thereby spread
// over all waves.  Ends with LDS free for the caller's next tile.  CHECKED
// false: the tile lies wholly inside M x N, no output is tested.
template <int WM, int WN, int TM, int TN, int BLK, int WK, bool CHECKED = true>
__device__ __forceinline__ void sgemm_finish(typename MfmaF32<BLK>::acc_t (&acc)[TM][TN], float* __restrict__ C,
                                             int M, int N, int64_t ldc, float alpha, float beta, int m0, int n0,
                                             int wm, int wn, int wmn, int wk, int lane, float* smem) {
    typedef MfmaF32<BLK> Op;
    constexpr int RS = Op::NREG / WK;
    static_assert(Op::NREG % WK == 0, "each wave group must own whole accumulator registers");
    if constexpr (WK > 1) {
        // partial sums of group g for owner o != g
        auto slot = [&](int g, int o, int i, int j, int r) {
            const int region = g * (WK - 1) + (o < g ? o : o - 1);
            return ((((region * WM * WN + wmn) * TM + i) * TN + j) * RS + r % RS) * 64 + lane;
        };
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int r = 0; r < Op::NREG; ++r)
                    if (r / RS != wk) smem[slot(wk, r / RS, i, j, r)] = acc[i][j][r];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int r = 0; r < Op::NREG; ++r)
                    if (r / RS == wk) {
                        float v = wk == 0 ? acc[i][j][r] : smem[slot(0, wk, i, j, r)];
#pragma unroll
                        for (int g = 1; g < WK; ++g) v += g == wk ? acc[i][j][r] : smem[slot(g, wk, i, j, r)];
                        acc[i][j][r] = v;
                    }
        __syncthreads();  // LDS free again before the caller's next tile
    }

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < Op::NREG; ++r) {
                if (WK > 1 && r / RS != wk) continue;
                const int row = m0 + wm * TM * BLK + i * BLK + Op::out_row(lane, r);
                const int col = n0 + wn * TN * BLK + j * BLK + Op::out_col(lane);
                if (!CHECKED || (row < M && col < N)) {
                    float* p = C + (int64_t)row * ldc + col;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * gload<float>(p);
                    gstore(p, v);
                }
            }
}

// Computes the BM x BN tile at (m0, n0) of C = alpha*op(A)op(B) + beta*C.
// Must be called by all WM*WN*WK*64 threads of the group; leaves LDS free
// (ends after a barrier).
template <int WM, int WN, int TM, int TN, int BK, int BLK, bool TA, bool TB, int WK = 1>
__device__ __forceinline__ void sgemm_tile(const float* __restrict__ A, const float* __restrict__ B,
                                           float* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb,
                                           int64_t ldc, float alpha, float beta, int m0, int n0, bool vec_ok,
                                           float* smem) {
    typedef MfmaF32<BLK> Op;
    typedef SgemmShape<WM, WN, TM, TN, BK, BLK, WK> S;
    constexpr int NT = S::NT, BM = S::BM, BN = S::BN, LDK = S::LDK;
    static_assert(BK % Op::KGROUP == 0, "BK must cover whole MFMA k-groups");
    constexpr int NSTEP = BK / Op::KGROUP;  // MFMA k-groups per staged K tile
    static_assert(NSTEP % WK == 0, "the k-split groups must share a stage's steps evenly");
    constexpr int NS = NSTEP / WK;  // steps per wave and stage
    typedef TileIO<BM, BK, NT, TA> IoA;   // A is direct when op(A) rows are k-contiguous
    typedef TileIO<BN, BK, NT, !TB> IoB;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wmn = wave % (WM * WN), wk = wave / (WM * WN);  // wk = 0 when WK == 1
    const int wm = wmn / WN, wn = wmn % WN;
    const bool full_mn = vec_ok && (m0 + BM <= M) && (n0 + BN <= N);

    typename Op::acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < Op::NREG; ++r) acc[i][j][r] = 0.0f;

    f4v ra[IoA::NV], rb[IoB::NV];
    const int KT = (K + BK - 1) / BK;
    auto fetch = [&](int t) {
        const int k0 = t * BK;
        const bool full = full_mn && (k0 + BK <= K);
        IoA::load(ra, A, lda, m0, k0, M, K, full, vec_ok, tid);
        IoB::load(rb, B, ldb, n0, k0, N, K, full, vec_ok, tid);
    };

    fetch(0);
    IoA::store(ra, smem, tid);
    IoB::store(rb, smem + BM * LDK, tid);
    __syncthreads();

    const int arow = wm * TM * BLK + Op::lane_row(lane);
    const int brow = wn * TN * BLK + Op::lane_row(lane);
    const int kq = Op::lane_kq(lane) + wk * NS * Op::KGROUP;  // this k-group's run of steps

    for (int t = 0; t < KT; ++t) {
        const float* As = smem + (t & 1) * (BM + BN) * LDK;
        const float* Bs = As + BM * LDK;
        if (t + 1 < KT) fetch(t + 1);
        // The next stage goes to LDS halfway through this one's steps, not
        // after the last: its stores, and the barrier after them, then
        // overlap the MFMAs of the second half still in the pipeline.
        constexpr int STORE_AT = NS / 2 - 1 < 0 ? 0 : NS / 2 - 1;
        // With B staged through transposing stores (B not transposed), let
        // the scheduler interleave the fragment reads with the MFMAs: NN at
        // 8192^3 135 -> 137.5 TFLOP/s; NT loses 0.6 % with it, so not there.
        if constexpr (!TB) __builtin_amdgcn_iglp_opt(0);
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            const int kk = st * Op::KGROUP;
            f4v a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = *(const f4v*)(As + (arow + i * BLK) * LDK + kk + kq);
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = *(const f4v*)(Bs + (brow + j * BLK) * LDK + kk + kq);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j) acc[i][j] = Op::mma(a[i][s], b[j][s], acc[i][j]);
            if (st == STORE_AT && t + 1 < KT) {
                float* nA = smem + ((t + 1) & 1) * (BM + BN) * LDK;
                IoA::store(ra, nA, tid);
                IoB::store(rb, nA + BM * LDK, tid);
            }
        }
        __syncthreads();
    }

    sgemm_finish<WM, WN, TM, TN, BLK, WK>(acc, C, M, N, ldc, alpha, beta, m0, n0, wm, wn, wmn, wk, lane, smem);
}

}  // namespace gat
