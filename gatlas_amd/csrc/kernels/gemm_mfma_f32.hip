// FP32 GEMM on gfx950 matrix cores: one work-group per macro tile, walked
// in XCD-aware supertile order; the tile computation is sgemm_tile
// (sgemm_tile.hpp), shared with the persistent grouped kernel.  The tiles
// staged 32 deep (32x32 blocks, no k-split or a two-group one) run their
// interior tiles (K >= 32 a multiple of 16, 16-byte aligned rows) on the
// direct-to-LDS routine (sgemm_dma_tile.hpp), which gives the same bits.
#include "sgemm_dma_tile.hpp"
#include "sgemm_tile.hpp"

namespace gat {

// Configurations whose interior tiles take the direct-to-LDS routine.
template <int WM, int WN, int TM, int TN, int BK, int BLK, int WK>
constexpr bool sgemm_has_dma() {
    constexpr int NW = WM * WN * WK, BM = WM * TM * 32, BN = WN * TN * 32;
    return (WK == 1 || WK == 2) && BK == 32 && BLK == 32 && BM >= 64 && BN >= 64 && BM <= 256 && BN <= 256 &&
           (BM / 8) % NW == 0 && (BN / 8) % NW == 0;
}

template <int WM, int WN, int TM, int TN, int BK, int BLK, int WK>
constexpr size_t sgemm_smem() {
    constexpr size_t reg = SgemmShape<WM, WN, TM, TN, BK, BLK, WK>::SMEM;
    constexpr size_t dma =
        sgemm_has_dma<WM, WN, TM, TN, BK, BLK, WK>() ? SgemmDmaShape<WM, WN, TM, TN, WK>::SMEM : 0;
    return reg > dma ? reg : dma;
}

template <int WM, int WN, int TM, int TN, int BK, int BLK, int WK, bool TA, bool TB>
__global__ __launch_bounds__(WM* WN * WK * 64) void sgemm_mfma_kernel(
    const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ C, int M, int N, int K,
    int64_t lda, int64_t ldb, int64_t ldc, int64_t sA, int64_t sB, int64_t sC, float alpha, float beta,
    int tiles_m, int tiles_n, int group_m, int vec_ok) {
    typedef SgemmShape<WM, WN, TM, TN, BK, BLK, WK> S;
    extern __shared__ float smem[];
    A += (int64_t)blockIdx.y * sA;
    B += (int64_t)blockIdx.y * sB;
    C += (int64_t)blockIdx.y * sC;
    int tmi, tni;
    mfma_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    if constexpr (sgemm_has_dma<WM, WN, TM, TN, BK, BLK, WK>()) {
        const int m0 = tmi * S::BM, n0 = tni * S::BN;
        if (vec_ok && K >= 32 && K % 16 == 0 && m0 + S::BM <= M && n0 + S::BN <= N) {
            sgemm_tile_dma<WM, WN, TM, TN, TA, TB, WK>(A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, m0, n0, smem);
            return;
        }
    }
    sgemm_tile<WM, WN, TM, TN, BK, BLK, TA, TB, WK>(A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, tmi * S::BM,
                                                    tni * S::BN, vec_ok != 0, smem);
}

// Instantiated configurations: (WM, WN, TM, TN, BK, BLK, WK).  The WK > 1
// entries split each staged K tile between WK wave groups: small problems,
// whose tiles barely cover the CUs, then keep several waves per SIMD.
#define GAT_F32_MFMA_CFGS(X)  \
    X(2, 2, 2, 2, 16, 32, 1)  \
    X(2, 2, 2, 2, 32, 32, 1)  \
    X(2, 2, 4, 2, 16, 32, 1)  \
    X(2, 2, 2, 4, 16, 32, 1)  \
    X(2, 2, 4, 4, 16, 32, 1)  \
    X(2, 2, 1, 1, 32, 32, 1)  \
    X(1, 1, 2, 2, 32, 32, 1)  \
    X(2, 2, 2, 2, 32, 16, 1)  \
    X(2, 2, 4, 4, 16, 16, 1)  \
    X(2, 2, 4, 4, 8, 32, 1)   \
    X(2, 4, 4, 2, 16, 32, 1)  \
    X(2, 4, 4, 2, 32, 32, 1)  \
    X(4, 2, 2, 4, 16, 32, 1)  \
    X(4, 4, 2, 2, 16, 32, 1)  \
    X(4, 4, 2, 2, 32, 32, 1)  \
    X(2, 2, 1, 1, 32, 32, 2)  \
    X(2, 2, 1, 1, 64, 32, 4)  \
    X(2, 2, 2, 2, 32, 32, 2)  \
    X(2, 2, 2, 2, 64, 32, 2)  \
    X(4, 4, 1, 1, 32, 32, 1)  \
    X(2, 4, 2, 1, 32, 32, 1)  \
    X(4, 2, 1, 2, 32, 32, 1)

static const MfmaCfg kF32Cfgs[] = {
#define X(wm, wn, tm, tn, bk, blk, wk) {wm, wn, tm, tn, bk, blk, wk},
    GAT_F32_MFMA_CFGS(X)
#undef X
};

int mfma_f32_num_cfgs() { return (int)(sizeof(kF32Cfgs) / sizeof(kF32Cfgs[0])); }
MfmaCfg mfma_f32_cfg(int i) { return kF32Cfgs[i]; }

bool mfma_f32_has(const MfmaCfg& c) {
    for (const MfmaCfg& k : kF32Cfgs)
        if (k.waves_m == c.waves_m && k.waves_n == c.waves_n && k.tm == c.tm && k.tn == c.tn &&
            k.bk == c.bk && k.blk == c.blk && k.wk == c.wk)
            return true;
    return false;
}

template <int WM, int WN, int TM, int TN, int BK, int BLK, int WK>
static int launch_one(const GemmArgs& a, hipStream_t st) {
    typedef SgemmShape<WM, WN, TM, TN, BK, BLK, WK> S;
    const int tiles_m = (int)((a.M + S::BM - 1) / S::BM), tiles_n = (int)((a.N + S::BN - 1) / S::BN);
    const float* A = (const float*)a.A;
    const float* B = (const float*)a.B;
    const int vec_ok = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && (a.lda % 4 == 0) &&
                       (a.ldb % 4 == 0) && (a.strideA % 4 == 0) && (a.strideB % 4 == 0);
    typedef void (*kfn)(const float*, const float*, float*, int, int, int, int64_t, int64_t, int64_t,
                        int64_t, int64_t, int64_t, float, float, int, int, int, int);
    kfn f;
    if (!a.transA && !a.transB) f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, WK, false, false>;
    else if (a.transA && !a.transB) f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, WK, true, false>;
    else if (!a.transA && a.transB) f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, WK, false, true>;
    else f = sgemm_mfma_kernel<WM, WN, TM, TN, BK, BLK, WK, true, true>;
    constexpr size_t smem = sgemm_smem<WM, WN, TM, TN, BK, BLK, WK>();
    if (smem > 65536) GAT_HIP_TRY(allow_dynamic_lds((const void*)f, smem));
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)a.batch);
    hipLaunchKernelGGL(f, grid, dim3(WM * WN * WK * 64), smem, st, A, B, (float*)a.C, (int)a.M, (int)a.N,
                       (int)a.K, a.lda, a.ldb, a.ldc, a.strideA, a.strideB, a.strideC, (float)a.alpha,
                       (float)a.beta, tiles_m, tiles_n, group_m, vec_ok);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_mfma_f32(const MfmaCfg& c, const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) return ST_INVALID;
#define X(p_wm, p_wn, p_tm, p_tn, p_bk, p_blk, p_wk)                                            \
    if (c.waves_m == p_wm && c.waves_n == p_wn && c.tm == p_tm && c.tn == p_tn && c.bk == p_bk &&  \
        c.blk == p_blk && c.wk == p_wk)                                                           \
        return launch_one<p_wm, p_wn, p_tm, p_tn, p_bk, p_blk, p_wk>(a, st);
    GAT_F32_MFMA_CFGS(X)
#undef X
    set_error("mfma f32: configuration not instantiated");
    return ST_UNSUPPORTED;
}

}  // namespace gat
