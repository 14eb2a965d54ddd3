// FP32 GEMM on gfx950 matrix cores (v_mfma_f32_32x32x2_f32 / 16x16x4_f32).
//
// Work decomposition (one work-group per macro tile, XCD-aware order):
//   * macro tile BM x BN = (WM*TM*BLK) x (WN*TN*BLK), WM x WN wavefronts,
//     each wave owns TM x TN output blocks of BLK x BLK held in accumulators;
//   * K is staged through double-buffered LDS, BK deep.  Both operands are
//     kept k-contiguous in LDS (A as [m][k], B as [n][k], row stride BK+4
//     floats) so one ds_read_b128 gives a lane the operands of four MFMAs;
//     the +4 pad spreads the 32 rows a half-wave reads over distinct banks;
//   * the next K tile is fetched from HBM into registers while the current
//     one is multiplied, then written to the other LDS buffer: one barrier
//     per K step, and no write-after-read hazard on the LDS tiles;
//   * epilogue C = alpha*acc + beta*C (C untouched when beta == 0), with
//     each half-wave storing 32 consecutive floats of a row.
// MFMA k-lane mapping: MFMA s of a group takes, in lane l, the k index
// kk + 4*(l/32) + s (32x32x2) or kk + 4*(l/16) + s (16x16x4).  A and B use
// the same mapping, so the product is the full sum over the group's k.
#include "../kernels.hpp"

namespace gat {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Work-group id -> (tile_m, tile_n).  Consecutive ids land on different XCDs
// (round-robin dispatch), so ids are first regrouped so that each XCD owns a
// contiguous range, then walked in column-major supertiles `group_m` tiles
// tall so that concurrently resident tiles share A and B panels in L2.
__device__ __forceinline__ void tile_coords(int wg, int nwg, int tiles_m, int tiles_n, int group_m,
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
    __device__ static __forceinline__ int out_row(int lane, int r) {
        return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
    }
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

    __device__ static __forceinline__ void load(f4v (&reg)[NV], const float* __restrict__ src, int64_t ld,
                                                int r0, int k0, int rows_lim, int k_lim, bool full,
                                                int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int idx = q * NT + tid;
            if (!TRANS) {
                const int r = idx / (BK / 4), kq = (idx % (BK / 4)) * 4;
                const int gr = r0 + r, gk = k0 + kq;
                if (full) {
                    reg[q] = *(const f4v*)(src + (int64_t)gr * ld + gk);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        reg[q][e] = (gr < rows_lim && gk + e < k_lim) ? src[(int64_t)gr * ld + gk + e] : 0.0f;
                }
            } else {
                const int k = idx % BK, rq = (idx / BK) * 4;
                const int gk = k0 + k, gr = r0 + rq;
                if (full) {
                    reg[q] = *(const f4v*)(src + (int64_t)gk * ld + gr);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        reg[q][e] = (gk < k_lim && gr + e < rows_lim) ? src[(int64_t)gk * ld + gr + e] : 0.0f;
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

template <int WM, int WN, int TM, int TN, int BK, int BLK, bool TA, bool TB>
__global__ __launch_bounds__(WM* WN * 64) void sgemm_mfma_kernel(
    const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ C, int M, int N, int K,
    int64_t lda, int64_t ldb, int64_t ldc, int64_t sA, int64_t sB, int64_t sC, float alpha, float beta,
    int tiles_m, int tiles_n, int group_m, int vec_ok) {
    typedef MfmaF32<BLK> Op;
    constexpr int NT = WM * WN * 64;
    constexpr int BM = WM * TM * BLK, BN = WN * TN * BLK;
    constexpr int LDK = BK + 4;
    static_assert(BK % Op::KGROUP == 0, "BK must cover whole MFMA k-groups");
    // A: direct when op(A) rows are k-contiguous (not transposed).
    typedef TileIO<BM, BK, NT, TA> IoA;
    typedef TileIO<BN, BK, NT, !TB> IoB;

    extern __shared__ float smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;

    A += (int64_t)blockIdx.y * sA;
    B += (int64_t)blockIdx.y * sB;
    C += (int64_t)blockIdx.y * sC;

    int tmi, tni;
    tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    const int m0 = tmi * BM, n0 = tni * BN;
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

    auto gload = [&](int t) {
        const int k0 = t * BK;
        const bool full = full_mn && (k0 + BK <= K);
        IoA::load(ra, A, lda, m0, k0, M, K, full, tid);
        IoB::load(rb, B, ldb, n0, k0, N, K, full, tid);
    };

    gload(0);
    IoA::store(ra, smem, tid);
    IoB::store(rb, smem + BM * LDK, tid);
    __syncthreads();

    const int arow = wm * TM * BLK + Op::lane_row(lane);
    const int brow = wn * TN * BLK + Op::lane_row(lane);
    const int kq = Op::lane_kq(lane);

    for (int t = 0; t < KT; ++t) {
        const float* As = smem + (t & 1) * (BM + BN) * LDK;
        const float* Bs = As + BM * LDK;
        if (t + 1 < KT) gload(t + 1);
#pragma unroll
        for (int kk = 0; kk < BK; kk += Op::KGROUP) {
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
        }
        if (t + 1 < KT) {
            float* nA = smem + ((t + 1) & 1) * (BM + BN) * LDK;
            IoA::store(ra, nA, tid);
            IoB::store(rb, nA + BM * LDK, tid);
        }
        __syncthreads();
    }

    // epilogue
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < Op::NREG; ++r) {
                const int row = m0 + wm * TM * BLK + i * BLK + Op::out_row(lane, r);
                const int col = n0 + wn * TN * BLK + j * BLK + Op::out_col(lane);
                if (row < M && col < N) {
                    float* p = C + (int64_t)row * ldc + col;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *p;
                    *p = v;
                }
            }
}

// Instantiated configurations: (WM, WN, TM, TN, BK, BLK).
#define GAT_F32_MFMA_CFGS(X) \
    X(2, 2, 2, 2, 16, 32)    \
    X(2, 2, 2, 2, 32, 32)    \
    X(2, 2, 4, 2, 16, 32)    \
    X(2, 2, 2, 4, 16, 32)    \
    X(2, 2, 4, 4, 16, 32)    \
    X(2, 2, 1, 1, 32, 32)    \
    X(1, 1, 2, 2, 32, 32)    \
    X(2, 2, 2, 2, 32, 16)    \
    X(2, 2, 4, 4, 16, 16)

static const MfmaCfg kF32Cfgs[] = {
#define X(wm, wn, tm, tn, bk, blk) {wm, wn, tm, tn, bk, blk},
    GAT_F32_MFMA_CFGS(X)
#undef X
};

int mfma_f32_num_cfgs() { return (int)(sizeof(kF32Cfgs) / sizeof(kF32Cfgs[0])); }
MfmaCfg mfma_f32_cfg(int i) { return kF32Cfgs[i]; }

bool mfma_f32_has(const MfmaCfg& c) {
    for (const MfmaCfg& k : kF32Cfgs)
        if (k.waves_m == c.waves_m && k.waves_n == c.waves_n && k.tm == c.tm && k.tn == c.tn &&
            k.bk == c.bk && k.blk == c.blk)
            return true;
    return false;
}

template <int WM, int WN, int TM, int TN, int BK, int BLK>
static int launch_one(const GemmArgs& a, hipStream_t st) {
    constexpr int BM = WM * TM * BLK, BN = WN * TN * BLK;
    constexpr size_t smem = 2 * (size_t)(BM + BN) * (BK + 4) * sizeof(float);
    const int tiles_m = (int)((a.M + BM - 1) / BM), tiles_n = (int)((a.N + BN - 1) / BN);
    const float* A = (const float*)a.A;
    const float* B = (const float*)a.B;
    const int vec_ok = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && (a.lda % 4 == 0) &&
                       (a.ldb % 4 == 0) && (a.strideA % 4 == 0) && (a.strideB % 4 == 0);
    typedef void (*kfn)(const float*, const float*, float*, int, int, int, int64_t, int64_t, int64_t,
                        int64_t, int64_t, int64_t, float, float, int, int, int, int);
    kfn f;
    if (!a.transA && !a.transB) f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, false, false>;
    else if (a.transA && !a.transB) f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, true, false>;
    else if (!a.transA && a.transB) f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, false, true>;
    else f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, true, true>;
    if (smem > 65536) GAT_HIP_TRY(hipFuncSetAttribute((const void*)f,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)a.batch);
    hipLaunchKernelGGL(f, grid, dim3(WM * WN * 64), smem, st, A, B, (float*)a.C, (int)a.M, (int)a.N,
                       (int)a.K, a.lda, a.ldb, a.ldc, a.strideA, a.strideB, a.strideC, (float)a.alpha,
                       (float)a.beta, tiles_m, tiles_n, group_m, vec_ok);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_mfma_f32(const MfmaCfg& c, const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) return ST_INVALID;
#define X(p_wm, p_wn, p_tm, p_tn, p_bk, p_blk)                                                  \
    if (c.waves_m == p_wm && c.waves_n == p_wn && c.tm == p_tm && c.tn == p_tn && c.bk == p_bk &&  \
        c.blk == p_blk)                                                                           \
        return launch_one<p_wm, p_wn, p_tm, p_tn, p_bk, p_blk>(a, st);
    GAT_F32_MFMA_CFGS(X)
#undef X
    set_error("mfma f32: configuration not instantiated");
    return ST_UNSUPPORTED;
}

}  // namespace gat
