// Split-BF16x3 GEMM through pre-split operand planes (gfx950).
//
// A bandwidth pass splits op(A) (M x K) and op(B)^T (N x K) -- each element
// a into hi = bf16(a) and lo = bf16(a - hi), the rounding of the in-kernel
// form (split_bf16x2) -- into bf16 planes laid out k-contiguous, every 32 k
// values stored as [32 hi | 32 lo]: 128 bytes per row per 32-deep K stage.
// Rows are padded with zeros to whole tiles and k to a multiple of 32, so
// every tile of the GEMM is interior.  The GEMM stages both planes straight
// into LDS (DmaOperand, k-contiguous, XOR-swizzled rows) and issues, per
// 16-deep step and block pair, lo*hi, hi*lo, hi*hi into FP32 accumulators:
// the products and the order of the register-staged split-BF16x3 tile, so
// the two forms agree bit for bit (zero padding adds exact zeros).
#include "bf16_dma_tile.hpp"

namespace gat {

constexpr int X3S_ROWS = 64;  // rows per work-group of the split pass
constexpr int X3S_K = 128;    // k values per work-group of the split pass

// Eight consecutive k values of one row split into the planes: hi words to
// d[0..3], lo words to d[16..19] (as dwords; the plane pair of a 32-k block
// is 32 hi then 32 lo bf16).
__device__ __forceinline__ void x3_split_store8(const float (&v)[8], __bf16* d) {
    unsigned h[4], l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) split_bf16x2(v[2 * j], v[2 * j + 1], h[j], l[j]);
    gstore(d, u32x4{h[0], h[1], h[2], h[3]});
    gstore(d + 32, u32x4{l[0], l[1], l[2], l[3]});
}

// dst row r (r < gridDim.y * 64): for each 32-k block b, 32 hi then 32 lo;
// src element (r, k) = src[r*rs + k*cs], zero outside rows x K.  A
// work-group splits 64 rows x 128 k values (k < kpad); work-item t takes
// rows t/16 + 16i (i < 4) and the eight k values 8*(t%16) .. +7 of each, so
// every 16 work-items write 512 contiguous bytes of a row's planes.
// k-contiguous sources (cs == 1) are read straight into registers, 16 bytes
// at a time where `vec`; row-contiguous ones (rs == 1) go through LDS,
// 16 bytes of four rows per load, and are read back a row at a time.
__global__ __launch_bounds__(256) void x3_split_kernel(const float* __restrict__ src, int64_t rs, int64_t cs,
                                                      int rows, int K, int kpad, int64_t ldd, int vec,
                                                      __bf16* __restrict__ dst) {
    __shared__ float tile[X3S_K][X3S_ROWS + 1];
    const int k0 = blockIdx.x * X3S_K, r0 = blockIdx.y * X3S_ROWS, tid = threadIdx.x;
    const int kk = k0 + 8 * (tid & 15);  // this work-item's first k value
    if (cs == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gr = r0 + (tid >> 4) + 16 * i;
            float v[8];
            const float* p = src + (int64_t)gr * rs + kk;
            if (vec && gr < rows && kk + 8 <= K) {
                const f4v a = gload<f4v>(p), b = gload<f4v>(p + 4);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    v[e] = a[e];
                    v[e + 4] = b[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (gr < rows && kk + e < K) ? gload<float>(p + e) : 0.0f;
            }
            if (kk < kpad) x3_split_store8(v, dst + (int64_t)gr * ldd + (kk >> 5) * 64 + (kk & 31));
        }
        return;
    }
    if (rs == 1) {
        // 128 k-rows of 64 rows: 16 float4 per k-row, 8 per work-item
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = (tid >> 4) + 16 * i, rq = 4 * (tid & 15);
            const int gk = k0 + k, gr = r0 + rq;
            const float* p = src + (int64_t)gk * cs + gr;
            if (vec && gk < K && gr + 4 <= rows) {
                const f4v a = gload<f4v>(p);
#pragma unroll
                for (int e = 0; e < 4; ++e) tile[k][rq + e] = a[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) tile[k][rq + e] = (gk < K && gr + e < rows) ? gload<float>(p + e) : 0.0f;
            }
        }
    } else {
        for (int q = tid; q < X3S_K * X3S_ROWS; q += 256) {
            const int k = q >> 6, r = q & 63, gk = k0 + k, gr = r0 + r;
            tile[k][r] = (gk < K && gr < rows) ? gload<float>(src + (int64_t)gr * rs + (int64_t)gk * cs) : 0.0f;
        }
    }
    __syncthreads();
    if (kk >= kpad) return;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = (tid >> 4) + 16 * i;
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = tile[kk - k0 + e][r];
        x3_split_store8(v, dst + (int64_t)(r0 + r) * ldd + (kk >> 5) * 64 + (kk & 31));
    }
}

// The GEMM on split planes: As is Mp x 2*kpad, Bs is Np x 2*kpad (row
// stride ld2 elements), Mp and Np whole tiles.  Stage t holds k values
// 32t .. 32t+31 of both planes.
template <int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(WM* WN * 64) __attribute__((amdgpu_waves_per_eu(2))) void x3_split_gemm_kernel(
    const __bf16* __restrict__ As, const __bf16* __restrict__ Bs, float* __restrict__ C, int M, int N, int kpad,
    int64_t ld2, int64_t ldc, float alpha, float beta, int tiles_m, int tiles_n, int group_m) {
    constexpr int NW = WM * WN, BM = WM * TM * 32, BN = WN * TN * 32;
    typedef DmaOperand<BM, true, NW> OpA;
    typedef DmaOperand<BN, true, NW> OpB;
    constexpr int STAGE = OpA::BYTES + OpB::BYTES;
    extern __shared__ __attribute__((aligned(16))) __bf16 xsm[];
    char* const sb = (char*)xsm;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = w / WN, wn = w % WN;
    int tmi, tni;
    mfma_tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tmi, tni);
    const int m0 = tmi * BM, n0 = tni * BN;

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    // one 16-deep step's operands: hi and lo planes of the wave's A and B blocks
    struct Frags {
        bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
    };
    Frags f0, f1;
    // 16-deep step s of a stage: hi chunk 2s + lane/32, lo chunk 4 more
    auto frags = [&](const char* st, int s, Frags& f) {
        const int c = 2 * s + (lane >> 5);
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int row = wm * TM * 32 + i * 32 + (lane & 31);
            const u32x4 h = OpA::chunk(st, row, c), l = OpA::chunk(st, row, c + 4);
            f.ah[i] = *(const bf16x8*)&h;
            f.al[i] = *(const bf16x8*)&l;
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int row = wn * TN * 32 + j * 32 + (lane & 31);
            const u32x4 h = OpB::chunk(st + OpA::BYTES, row, c), l = OpB::chunk(st + OpA::BYTES, row, c + 4);
            f.bh[j] = *(const bf16x8*)&h;
            f.bl[j] = *(const bf16x8*)&l;
        }
    };
    auto mma3 = [&](const Frags& f) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.al[i], f.bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bl[j], acc[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bh[j], acc[i][j], 0, 0, 0);
    };
    // one MFMA, then one LDS read and one global load if any remain, ...
    auto interleave = [&]() {
#pragma unroll
        for (int i = 0; i < 3 * TM * TN; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load (to LDS)
        }
    };

    const int KT = kpad / 32;
    constexpr int LOADS = OpA::PER_WAVE + OpB::PER_WAVE;  // loads per wave per stage
    auto issue = [&](int t, char* buf) {
        OpA::issue(As, ld2, m0, 64 * t, buf, w, lane);
        OpB::issue(Bs, ld2, n0, 64 * t, buf + OpA::BYTES, w, lane);
    };
    issue(0, sb);
    if (KT > 1) {
        issue(1, sb + STAGE);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(LOADS, 0));  // stage 0 in LDS
    } else {
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));
    }
    __builtin_amdgcn_s_barrier();
    frags(sb, 0, f0);
    // The fragments of step 1 are read under the MFMAs of step 0.  Then each
    // wave waits for the next stage and its own reads of this one, and the
    // waves meet at a barrier; this stage's buffer is free, and the loads of
    // the stage after next go into it while step 1's MFMAs run under the
    // reads of the next stage's step 0, so a stage's loads have a whole
    // iteration to land.  Past the last stage the loads repeat it (unused).
    for (int t = 0; t < KT; ++t) {
        char* cur = sb + (t & 1) * STAGE;
        const char* nxt = sb + ((t + 1) & 1) * STAGE;
        __builtin_amdgcn_sched_barrier(0);
        frags(cur, 1, f1);
        mma3(f0);
        interleave();
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the next stage in; this wave's reads of this one done
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t + 2 < KT ? t + 2 : KT - 1, cur);
        frags(nxt, 0, f0);
        mma3(f1);
        interleave();
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_imm(0, 0));  // the repeated loads land before the wave ends

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * TM * 32 + i * 32 + mfma32_row(lane, r);
                const int col = n0 + wn * TN * 32 + j * 32 + mfma32_col(lane);
                if (row < M && col < N) {
                    float* p = C + (int64_t)row * ldc + col;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * gload<float>(p);
                    gstore(p, v);
                }
            }
}

// The one split-plane configuration: 256 x 256 tiles of 8 waves (4 x 2
// blocks each), 32 k values per stage; listed in the split-BF16x3 tile
// table with bk = 32.
static const MfmaCfg kX3Split{2, 4, 4, 2, 32, 32};

bool x3_split_cfg(const MfmaCfg& c) {
    return c.waves_m == kX3Split.waves_m && c.waves_n == kX3Split.waves_n && c.tm == kX3Split.tm &&
           c.tn == kX3Split.tn && c.bk == kX3Split.bk && c.blk == kX3Split.blk && c.wk == 1;
}
MfmaCfg x3_split_cfg_value() { return kX3Split; }

X3SplitDims x3_split_dims(int64_t M, int64_t N, int64_t K) {
    constexpr int BM = 256, BN = 256;
    X3SplitDims d;
    d.mpad = (M + BM - 1) / BM * BM;
    d.npad = (N + BN - 1) / BN * BN;
    d.kpad = (K + 31) / 32 * 32;
    d.bytes_a = (size_t)d.mpad * 2 * d.kpad * sizeof(__bf16);
    d.bytes_b = (size_t)d.npad * 2 * d.kpad * sizeof(__bf16);
    return d;
}

int launch_x3_split(const float* src, int64_t rs, int64_t cs, int rows, int K, int64_t rows_pad, int64_t kpad,
                    void* dst, hipStream_t st) {
    if (rows_pad % X3S_ROWS || kpad % 32 || kpad > INT32_MAX || rows_pad / X3S_ROWS > 65535) {
        set_error("split-bf16x3: padded shape out of range");
        return ST_INVALID;
    }
    // 16-byte loads along the contiguous direction: aligned base and stride
    const int64_t stride = cs == 1 ? rs : cs;
    const int vec = ((uintptr_t)src % 16 == 0) && (stride % 4 == 0);
    hipLaunchKernelGGL(x3_split_kernel, dim3((unsigned)((kpad + X3S_K - 1) / X3S_K), (unsigned)(rows_pad / X3S_ROWS)),
                       dim3(256), 0, st, src, rs, cs, rows, K, (int)kpad, 2 * kpad, vec, (__bf16*)dst);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

int launch_x3_split_gemm(const X3SplitDims& d, const void* As, const void* Bs, float* C, int M, int N,
                         int64_t ldc, float alpha, float beta, hipStream_t st) {
    constexpr int WM = 2, WN = 4, TM = 4, TN = 2, BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr size_t smem = 2 * (size_t)(BM + BN) * 128;
    if (d.kpad <= 0 || d.kpad > INT32_MAX) {
        set_error("split-bf16x3: K out of range");
        return ST_INVALID;
    }
    const int tiles_m = (int)(d.mpad / BM), tiles_n = (int)(d.npad / BN);
    const void* f = (const void*)x3_split_gemm_kernel<WM, WN, TM, TN>;
    GAT_HIP_TRY(allow_dynamic_lds(f, smem));
    const int group_m = tiles_m < 8 ? tiles_m : 8;
    hipLaunchKernelGGL((x3_split_gemm_kernel<WM, WN, TM, TN>), dim3((unsigned)(tiles_m * tiles_n)),
                       dim3(WM * WN * 64), smem, st, (const __bf16*)As, (const __bf16*)Bs, C, M, N, (int)d.kpad,
                       2 * d.kpad, ldc, alpha, beta, tiles_m, tiles_n, group_m);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
