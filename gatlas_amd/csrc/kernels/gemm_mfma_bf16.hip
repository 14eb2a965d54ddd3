// BF16 matrix-core GEMMs on gfx950: one work-group per macro tile in
// XCD-aware supertile order; the tile computation is bf16_tile
// (bf16_tile.hpp), shared with the persistent grouped kernel.  Interior
// tiles stage straight into LDS where a direct-to-LDS routine exists: BF16
// inputs at BK = 64 (bf16_dma_tile.hpp), split-BF16x3 at any BK
// (x3_dma_tile.hpp, 32-deep stages); both give the register path's bits
// (BF16: for K a multiple of 16; for K = 8 mod 16 the two paths pad the
// short stage at different ends and so group the k values differently).
#include "bf16_dma_tile.hpp"
#include "bf16_tile.hpp"
#include "x3_dma_tile.hpp"

namespace gat {

// LDS of a configuration: its register-staged stages or its direct-to-LDS
// stages, whichever is larger.
template <int WM, int WN, int TM, int TN, int BK, int MODE>
constexpr size_t bf16_smem() {
    constexpr size_t reg = Bf16Shape<WM, WN, TM, TN, BK, MODE>::SMEM;
    constexpr size_t dma = MODE == MODE_BF16 ? (BK == 64 ? Bf16DmaShape<WM, WN, TM, TN>::SMEM : 0)
                           : x3_has_dma<WM, WN, TM, TN>() ? X3DmaShape<WM, WN, TM, TN>::SMEM
                                                           : 0;
    return reg > dma ? reg : dma;
}

// Waves per SIMD the kernel asks the register allocator to allow: what the
// LDS footprint permits (160 KiB per CU), capped by what the 512 registers
// of a SIMD lane hold at 16 accumulators per 32x32 block plus a working set
// (fragments, staging, addresses; two operand planes in the x3 form).
// Without it the allocator may place the accumulators in AGPRs next to a
// full VGPR set and leave one wave per SIMD where LDS allows two.
template <int WM, int WN, int TM, int TN, int BK, int MODE>
constexpr int bf16_waves_per_simd() {
    constexpr int groups = (int)(163840 / bf16_smem<WM, WN, TM, TN, BK, MODE>());
    constexpr int by_lds = groups * WM * WN / 4;
    constexpr int by_regs = 512 / (TM * TN * 16 + (MODE == MODE_X3 ? 160 : 96));
    constexpr int w = by_lds < by_regs ? by_lds : by_regs;
    constexpr int need = (WM * WN + 3) / 4;  // the group's own waves per SIMD
    return w < need ? need : (w > 8 ? 8 : w);
}

template <int WM, int WN, int TM, int TN, int BK, int MODE, bool TA, bool TB>
__global__ __launch_bounds__(WM* WN * 64)
__attribute__((amdgpu_waves_per_eu(bf16_waves_per_simd<WM, WN, TM, TN, BK, MODE>()))) void gemm_bf16_kernel(
    const void* __restrict__ Av, const void* __restrict__ Bv, float* __restrict__ C, int M, int N, int K,
    int64_t lda, int64_t ldb, int64_t ldc, int64_t sA, int64_t sB, int64_t sC, float alpha, float beta,
    int tiles_m, int tiles_n, int group_m, int vec_ok) {
    typedef Bf16Shape<WM, WN, TM, TN, BK, MODE> S;
    constexpr size_t ES = MODE == MODE_X3 ? sizeof(float) : sizeof(__bf16);
    extern __shared__ __bf16 smemh[];
    const char* A = (const char*)Av + (int64_t)blockIdx.y * sA * ES;
    const char* B = (const char*)Bv + (int64_t)blockIdx.y * sB * ES;
    C += (int64_t)blockIdx.y * sC;
    int tmi, tni;
    mfma_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    if constexpr (MODE == MODE_BF16 && BK == 64) {
        // interior tile, K >= 64 a multiple of 8, 16-byte aligned rows: direct-to-LDS staging
        const int m0 = tmi * S::BM, n0 = tni * S::BN;
        if (vec_ok && K >= 64 && K % 8 == 0 && m0 + S::BM <= M && n0 + S::BN <= N) {
            bf16_tile_dma<WM, WN, TM, TN, TA, TB>((const __bf16*)A, (const __bf16*)B, C, M, N, K, lda, ldb, ldc,
                                                  alpha, beta, m0, n0, smemh);
            return;
        }
    }
    if constexpr (MODE == MODE_X3 && x3_has_dma<WM, WN, TM, TN>()) {
        // interior tile, K >= 32 a multiple of 16, 16-byte aligned rows
        const int m0 = tmi * S::BM, n0 = tni * S::BN;
        if (vec_ok && K >= 32 && K % 16 == 0 && m0 + S::BM <= M && n0 + S::BN <= N) {
            x3_tile_dma<WM, WN, TM, TN, TA, TB>((const float*)A, (const float*)B, C, M, N, K, lda, ldb, ldc, alpha,
                                                beta, m0, n0, (float*)smemh);
            return;
        }
    }
    bf16_tile<WM, WN, TM, TN, BK, MODE, TA, TB>(A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, tmi * S::BM,
                                               tni * S::BN, vec_ok != 0, smemh);
}

// Instantiated configurations (WM, WN, TM, TN, BK, forms); blocks are
// 32x32; forms: 1 = BF16 inputs, 2 = split-BF16x3, 3 = both.  The 8-wave
// 256 x 256 tiles keep 0.75 KiB of LDS reads per MFMA (4 x 2 blocks per
// wave); in the x3 form they stage two planes, so only BK = 16 fits.
#define GAT_BF16_CFGS(X)  \
    X(2, 2, 2, 2, 32, 3)  \
    X(2, 2, 2, 2, 64, 3)  \
    X(2, 2, 4, 2, 32, 3)  \
    X(2, 2, 2, 4, 32, 3)  \
    X(2, 2, 1, 1, 64, 3)  \
    X(1, 1, 2, 2, 64, 3)  \
    X(2, 4, 4, 2, 32, 1)  \
    X(2, 4, 4, 2, 64, 1)  \
    X(4, 4, 2, 2, 64, 1)  \
    X(2, 4, 4, 2, 16, 2)  \
    X(2, 2, 2, 2, 16, 2)  \
    X(2, 2, 1, 1, 32, 2)

struct Bf16CfgEntry {
    MfmaCfg cfg;
    int forms;
};
static const Bf16CfgEntry kBf16Cfgs[] = {
#define X(wm, wn, tm, tn, bk, forms) {{wm, wn, tm, tn, bk, 32}, forms},
    GAT_BF16_CFGS(X)
#undef X
};

static int form_bit(int dtype) { return dtype == DT_BF16X3 ? 2 : 1; }

// The split-BF16x3 table ends with the split-plane configuration
// (gemm_x3_split.hip), which runs through the context (it needs a workspace).
int mfma_bf16_num_cfgs(int dtype) {
    int n = 0;
    for (const Bf16CfgEntry& e : kBf16Cfgs) n += (e.forms & form_bit(dtype)) != 0;
    return n + (dtype == DT_BF16X3 ? 1 : 0);
}
MfmaCfg mfma_bf16_cfg(int dtype, int i) {
    for (const Bf16CfgEntry& e : kBf16Cfgs)
        if ((e.forms & form_bit(dtype)) && i-- == 0) return e.cfg;
    if (dtype == DT_BF16X3 && i == 0) return x3_split_cfg_value();
    return MfmaCfg{0, 0, 0, 0, 0, 0};
}
bool mfma_bf16_has(int dtype, const MfmaCfg& c) {
    if (dtype == DT_BF16X3 && x3_split_cfg(c)) return true;
    for (const Bf16CfgEntry& e : kBf16Cfgs) {
        const MfmaCfg& k = e.cfg;
        if ((e.forms & form_bit(dtype)) && k.waves_m == c.waves_m && k.waves_n == c.waves_n && k.tm == c.tm &&
            k.tn == c.tn && k.bk == c.bk && k.blk == c.blk)
            return true;
    }
    return false;
}

template <int WM, int WN, int TM, int TN, int BK, int MODE>
static int launch_bf16_one(const GemmArgs& a, hipStream_t st) {
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr size_t smem = bf16_smem<WM, WN, TM, TN, BK, MODE>();
    const int tiles_m = (int)((a.M + BM - 1) / BM), tiles_n = (int)((a.N + BN - 1) / BN);
    const int W = MODE == MODE_X3 ? 4 : 8;  // elements per 16-byte vector
    const int vec_ok = ((uintptr_t)a.A % 16 == 0) && ((uintptr_t)a.B % 16 == 0) && (a.lda % W == 0) &&
                       (a.ldb % W == 0) && (a.strideA % W == 0) && (a.strideB % W == 0);
    typedef void (*kfn)(const void*, const void*, float*, int, int, int, int64_t, int64_t, int64_t, int64_t, int64_t,
                        int64_t, float, float, int, int, int, int);
    kfn f;
    if (!a.transA && !a.transB) f = gemm_bf16_kernel<WM, WN, TM, TN, BK, MODE, false, false>;
    else if (a.transA && !a.transB) f = gemm_bf16_kernel<WM, WN, TM, TN, BK, MODE, true, false>;
    else if (!a.transA && a.transB) f = gemm_bf16_kernel<WM, WN, TM, TN, BK, MODE, false, true>;
    else f = gemm_bf16_kernel<WM, WN, TM, TN, BK, MODE, true, true>;
    if (smem > 65536) GAT_HIP_TRY(allow_dynamic_lds((const void*)f, smem));
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)a.batch);
    hipLaunchKernelGGL(f, grid, dim3(WM * WN * 64), smem, st, a.A, a.B, (float*)a.C, (int)a.M, (int)a.N, (int)a.K,
                       a.lda, a.ldb, a.ldc, a.strideA, a.strideB, a.strideC, (float)a.alpha, (float)a.beta, tiles_m,
                       tiles_n, group_m, vec_ok);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_gemm_mfma_bf16(int dtype, const MfmaCfg& c, const GemmArgs& a, hipStream_t st) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (c.wk != 1) {
        set_error("k-split wave groups exist for FP32 tiles only");
        return ST_UNSUPPORTED;
    }
    if (a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) return ST_INVALID;
    if (dtype != DT_BF16 && dtype != DT_BF16X3) return ST_INVALID;
    if (dtype == DT_BF16X3 && x3_split_cfg(c)) {
        set_error("mfma bf16: the split-plane configuration runs through the context");
        return ST_INVALID;
    }
    if (!mfma_bf16_has(dtype, c)) {
        set_error("mfma bf16: configuration not instantiated for this form");
        return ST_UNSUPPORTED;
    }
#define X(p_wm, p_wn, p_tm, p_tn, p_bk, p_forms)                                                    \
    if (c.waves_m == p_wm && c.waves_n == p_wn && c.tm == p_tm && c.tn == p_tn && c.bk == p_bk) {   \
        if constexpr ((p_forms & 1) != 0)                                                           \
            if (dtype == DT_BF16) return launch_bf16_one<p_wm, p_wn, p_tm, p_tn, p_bk, MODE_BF16>(a, st); \
        if constexpr ((p_forms & 2) != 0)                                                           \
            if (dtype == DT_BF16X3) return launch_bf16_one<p_wm, p_wn, p_tm, p_tn, p_bk, MODE_X3>(a, st); \
    }
    GAT_BF16_CFGS(X)
#undef X
    set_error("mfma bf16: configuration not instantiated");
    return ST_UNSUPPORTED;
}

// FP32 -> BF16 (round to nearest even), n elements.
__global__ void f32_to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = (__bf16)x[i];
}

int launch_convert_f32_bf16(const float* x, void* y, size_t n, hipStream_t st) {
    if (!n) return ST_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, (__bf16*)y, n);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
