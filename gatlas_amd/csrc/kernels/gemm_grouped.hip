// Persistent grouped GEMM: many independent problems of different shapes
// and transposes in one launch.
//
// The host splits the (problem, m0, n0) tiles into one queue per XCD: each
// problem, or each block of tile rows and columns of a large one, goes whole to one queue
// (longest first onto the least-loaded queue), so its operand panels are
// read into one XCD's L2; inside a queue the longest tiles start first.
// A grid of about two work-groups per CU loops: one work-item takes the
// next index of its XCD's queue from that queue's counter (atomicAdd), and
// once the home queue is drained, from the other queues in turn; the group
// computes the tile with the routine of the problem's precision, shape and
// transpose pair, and leaves when every queue is drained.
// counter[0..7] are the queue cursors and counter[8] counts the groups that
// have left; the last group to leave (every other group has made its final
// fetches by then) sets all of them back to zero, so consecutive launches --
// and replays of a captured graph holding this one kernel node -- start
// from zeroed counters.  They must be zero before the first launch.
//
// Tile shapes (grouped.hpp): 0 = 128 x 128 (2 x 2 waves of 2 x 2 32x32
// blocks), 1 = 64 x 64 (one block per wave), 2 = 128 x 64 and 3 = 64 x 128
// (2 x 1 and 1 x 2 blocks per wave), for edge strips of the FP32 and
// BF16 forms 4 = 128 x 32 and 5 = 32 x 128 (4 x 1 and 1 x 4 waves of one
// block, staged through registers), and for the BF16 form the merged
// 6 = 256 x 128 and 7 = 128 x 256 (4 x 2 and 2 x 4 blocks per wave, 32-deep
// stages straight into LDS); each work entry carries its tile's.
// Tiles of 64 and more on each side of problems with K >= 32 (BF16: 64) a
// multiple of 16, 16-byte aligned rows and whole 16-byte chunks across each
// row-contiguous operand stage straight into LDS, edge tiles included (rows
// past the matrix are loaded as copies of its last rows and not stored;
// bf16_dma_tile.hpp, sgemm_dma_tile.hpp, x3_dma_tile.hpp); the others stage
// through registers.  The split-BF16x3 form splits FP32 operands into
// bf16 hi/lo as each wave reads its fragments (direct-to-LDS tiles) or while
// a stage is stored to LDS (bf16_tile.hpp, MODE_X3, staged 16 deep for the
// 128-wide shapes, 32 for 64 x 64); two groups fit a CU.  Neither the
// staging path nor its depth changes an x3 tile's bits (the same split, and
// its products run in 16-deep steps in k order), so a problem's result
// equals a single launch's.
#include "bf16_dma_tile.hpp"
#include "bf16_tile.hpp"
#include "grouped.hpp"
#include "sgemm_dma_tile.hpp"
#include "sgemm_tile.hpp"
#include "x3_dma_tile.hpp"

namespace gat {

// LDS slots of the merged BF16 tiles' 32-deep stages (two: 48 KiB for the
// 256 x 128 tile; three, 72 KiB, ran 6 % slower on the 512-problem workload)
constexpr int GP_MERGED_SLOTS = 2;

// Whether a problem's tiles stage straight into LDS (edge tiles too: their
// rows past the matrix are loaded as copies and not stored): 16-byte
// aligned rows, K >= kmin a multiple of 16, and a whole number of 16-byte
// chunks (e elements) across each row-contiguous operand.
template <bool TA, bool TB>
__device__ __forceinline__ bool gp_dma_ok(const GroupedProblemDev& p, int kmin, int e) {
    return (p.flags & GP_VEC) != 0 && p.K >= kmin && p.K % 16 == 0 && (!TA || p.M % e == 0) &&
           (TB || p.N % e == 0);
}

// FP32 tiles of 64 and more on each side stage straight into LDS where
// gp_dma_ok holds (sgemm_dma_tile.hpp, the same bits as the register path).
template <int WM, int WN, int TM, int TN, bool TA, bool TB>
__device__ __forceinline__ void f32_tile_t(const GroupedProblemDev& p, int m0, int n0, float* smem) {
    const bool vec = (p.flags & GP_VEC) != 0;
    const float* A = (const float*)p.A;
    const float* B = (const float*)p.B;
    if constexpr (WM * TM >= 2 && WN * TN >= 2) {
        if (gp_dma_ok<TA, TB>(p, 32, 4)) {
            sgemm_tile_dma<WM, WN, TM, TN, TA, TB, 1, true>(A, B, p.C, p.M, p.N, p.K, p.lda, p.ldb, p.ldc, p.alpha,
                                                            p.beta, m0, n0, smem);
            return;
        }
    }
    sgemm_tile<WM, WN, TM, TN, 32, 32, TA, TB>(A, B, p.C, p.M, p.N, p.K, p.lda, p.ldb, p.ldc, p.alpha, p.beta, m0,
                                               n0, vec, smem);
}

template <bool TA, bool TB>
__device__ __forceinline__ void f32_tile_shape(const GroupedProblemDev& p, int shape, int m0, int n0, float* smem) {
    if (shape == 0) f32_tile_t<2, 2, 2, 2, TA, TB>(p, m0, n0, smem);
    else if (shape == 1) f32_tile_t<2, 2, 1, 1, TA, TB>(p, m0, n0, smem);
    else if (shape == 2) f32_tile_t<2, 2, 2, 1, TA, TB>(p, m0, n0, smem);
    else if (shape == 3) f32_tile_t<2, 2, 1, 2, TA, TB>(p, m0, n0, smem);
    else if (shape == 4) f32_tile_t<4, 1, 1, 1, TA, TB>(p, m0, n0, smem);
    else f32_tile_t<1, 4, 1, 1, TA, TB>(p, m0, n0, smem);
}

// BF16 tiles of 64 and more on each side stage straight into LDS where
// gp_dma_ok holds (K >= 64; with K a multiple of 16 the routine's 16-deep
// steps are the register path's, so the bits do not depend on the path);
// the others through registers, BK deep.
template <int WM, int WN, int TM, int TN, int BK, bool TA, bool TB>
__device__ __forceinline__ void bf16_tile_t(const GroupedProblemDev& p, int m0, int n0, __bf16* smem) {
    const bool vec = (p.flags & GP_VEC) != 0;
    if constexpr (WM * TM >= 2 && WN * TN >= 2) {
        if (gp_dma_ok<TA, TB>(p, 64, 8)) {
            bf16_tile_dma<WM, WN, TM, TN, TA, TB, true>((const __bf16*)p.A, (const __bf16*)p.B, p.C, p.M, p.N, p.K,
                                                        p.lda, p.ldb, p.ldc, p.alpha, p.beta, m0, n0, smem);
            return;
        }
    }
    bf16_tile<WM, WN, TM, TN, BK, MODE_BF16, TA, TB>(p.A, p.B, p.C, p.M, p.N, p.K, p.lda, p.ldb, p.ldc, p.alpha,
                                                     p.beta, m0, n0, vec, smem);
}

// Merged BF16 tile (shape 6: TM = 4, 256 x 128; shape 7: TN = 4, 128 x 256):
// 32-deep stages straight into LDS in GP_MERGED_SLOTS slots
// (bf16_dma_tile.hpp).  The plan makes one only where that routine applies;
// anywhere else it is the two 128 x 128 tiles it stands for, which give the
// same bits.
template <int TM, int TN, bool TA, bool TB>
__device__ __forceinline__ void bf16_tile_merged(const GroupedProblemDev& p, int m0, int n0, __bf16* smem) {
    constexpr int BM = 2 * TM * 32, BN = 2 * TN * 32;
    const bool vec = (p.flags & GP_VEC) != 0;
    if (vec && p.K >= 32 && p.K % 16 == 0 && m0 + BM <= p.M && n0 + BN <= p.N) {
        bf16_tile_dma_ring<2, 2, TM, TN, GP_MERGED_SLOTS, TA, TB>((const __bf16*)p.A, (const __bf16*)p.B, p.C,
                                                                  p.M, p.N, p.K, p.lda, p.ldb, p.ldc, p.alpha,
                                                                  p.beta, m0, n0, smem);
        return;
    }
    bf16_tile_t<2, 2, 2, 2, 32, TA, TB>(p, m0, n0, smem);
    __syncthreads();  // every wave is done with the LDS before the second tile
    bf16_tile_t<2, 2, 2, 2, 32, TA, TB>(p, TM == 4 ? m0 + 128 : m0, TM == 4 ? n0 : n0 + 128, smem);
}

template <bool TA, bool TB>
__device__ __forceinline__ void bf16_tile_shape(const GroupedProblemDev& p, int shape, int m0, int n0,
                                                __bf16* smem) {
    if (shape == 0) bf16_tile_t<2, 2, 2, 2, 32, TA, TB>(p, m0, n0, smem);
    else if (shape == 1) bf16_tile_t<2, 2, 1, 1, 64, TA, TB>(p, m0, n0, smem);
    else if (shape == 2) bf16_tile_t<2, 2, 2, 1, 32, TA, TB>(p, m0, n0, smem);
    else if (shape == 3) bf16_tile_t<2, 2, 1, 2, 32, TA, TB>(p, m0, n0, smem);
    else if (shape == 4) bf16_tile_t<4, 1, 1, 1, 64, TA, TB>(p, m0, n0, smem);
    else if (shape == 5) bf16_tile_t<1, 4, 1, 1, 64, TA, TB>(p, m0, n0, smem);
    else if (shape == 6) bf16_tile_merged<4, 2, TA, TB>(p, m0, n0, smem);
    else bf16_tile_merged<2, 4, TA, TB>(p, m0, n0, smem);
}

// Split-BF16x3 tiles stage the FP32 operands straight into LDS where
// gp_dma_ok holds and split them as the fragments are read (x3_dma_tile.hpp,
// the same bits); the others split while staging through registers, BK
// deep.
template <int WM, int WN, int TM, int TN, int BK, bool TA, bool TB>
__device__ __forceinline__ void x3_tile_t(const GroupedProblemDev& p, int m0, int n0, __bf16* smem) {
    const bool vec = (p.flags & GP_VEC) != 0;
    if constexpr (x3_has_dma<WM, WN, TM, TN>()) {
        if (gp_dma_ok<TA, TB>(p, 32, 4)) {
            x3_tile_dma<WM, WN, TM, TN, TA, TB, true>((const float*)p.A, (const float*)p.B, p.C, p.M, p.N, p.K,
                                                      p.lda, p.ldb, p.ldc, p.alpha, p.beta, m0, n0, (float*)smem);
            return;
        }
    }
    bf16_tile<WM, WN, TM, TN, BK, MODE_X3, TA, TB>(p.A, p.B, p.C, p.M, p.N, p.K, p.lda, p.ldb, p.ldc, p.alpha,
                                                   p.beta, m0, n0, vec, smem);
}

template <bool TA, bool TB>
__device__ __forceinline__ void x3_tile_shape(const GroupedProblemDev& p, int shape, int m0, int n0,
                                              __bf16* smem) {
    if (shape == 0) x3_tile_t<2, 2, 2, 2, 16, TA, TB>(p, m0, n0, smem);
    else if (shape == 1) x3_tile_t<2, 2, 1, 1, 32, TA, TB>(p, m0, n0, smem);
    else if (shape == 2) x3_tile_t<2, 2, 2, 1, 16, TA, TB>(p, m0, n0, smem);
    else x3_tile_t<2, 2, 1, 2, 16, TA, TB>(p, m0, n0, smem);  // no 32-wide edge strips (gp_edge32)
}

// LDS of the tile buffers (the largest shape: 128 x 128; the 128 x 64
// shapes take less at the same depth).  The kernel
// declares no static LDS, so the dynamic block -- and every 16-byte tile
// access in it -- starts aligned.
constexpr size_t gp_max(size_t a, size_t b) { return a > b ? a : b; }

template <int FORM>
constexpr size_t grouped_tile_bytes() {
    // register-staged stages of shapes 0, 1, 2 and 4 (3 and 5 take what 2
    // and 4 do; the x3 form has no shapes 4 and 5)
    constexpr size_t reg =
        FORM == GP_FORM_F32
            ? gp_max(gp_max(SgemmShape<2, 2, 2, 2, 32, 32>::SMEM, SgemmShape<2, 2, 1, 1, 32, 32>::SMEM),
                     gp_max(SgemmShape<2, 2, 2, 1, 32, 32>::SMEM, SgemmShape<4, 1, 1, 1, 32, 32>::SMEM))
        : FORM == GP_FORM_BF16
            ? gp_max(gp_max(Bf16Shape<2, 2, 2, 2, 32, MODE_BF16>::SMEM, Bf16Shape<2, 2, 1, 1, 64, MODE_BF16>::SMEM),
                     gp_max(Bf16Shape<2, 2, 2, 1, 32, MODE_BF16>::SMEM, Bf16Shape<4, 1, 1, 1, 64, MODE_BF16>::SMEM))
            : gp_max(gp_max(Bf16Shape<2, 2, 2, 2, 16, MODE_X3>::SMEM, Bf16Shape<2, 2, 1, 1, 32, MODE_X3>::SMEM),
                     Bf16Shape<2, 2, 2, 1, 16, MODE_X3>::SMEM);
    // direct-to-LDS stages (the 128 x 128 shape's are the largest)
    constexpr size_t dma = FORM == GP_FORM_BF16 ? gp_max(Bf16DmaShape<2, 2, 2, 2>::SMEM,
                                                         Bf16RingShape<2, 2, 4, 2, GP_MERGED_SLOTS>::SMEM)
                           : FORM == GP_FORM_F32 ? SgemmDmaShape<2, 2, 2, 2>::SMEM
                                                 : X3DmaShape<2, 2, 2, 2>::SMEM;
    return (gp_max(reg, dma) + 15) / 16 * 16;
}

// Work-groups per CU, as LDS allows (160 KiB per CU, tile buffers plus the
// 16-byte work-index slot of each group), at most three.
template <int FORM>
constexpr int grouped_wgs_per_cu() {
    constexpr int n = (int)(163840 / (grouped_tile_bytes<FORM>() + 16));
    return n > 3 ? 3 : n;
}

// The register allocator is asked for as many waves per SIMD as work-groups
// fit a CU (a group has one wave per SIMD), which the union of all tile
// routines would otherwise exceed.
template <int FORM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grouped_wgs_per_cu<FORM>()))) void grouped_gemm_kernel(const GroupedProblemDev* __restrict__ probs,
                                                           const int4* __restrict__ work, GroupedQueues q,
                                                           int* __restrict__ counter) {
    extern __shared__ __attribute__((aligned(16))) float gsmem[];
    int& next = *(int*)((char*)gsmem + grouped_tile_bytes<FORM>());
    const int home = (int)(blockIdx.x % GP_XCDS);
    int qi = 0;  // queues tried so far: the home queue first, then the others in turn
    for (;;) {
        __syncthreads();  // every work-item has read the previous `next`
        if (threadIdx.x == 0) {
            int w = -1;
            for (; qi < GP_XCDS; ++qi) {
                const int x = (home + qi) % GP_XCDS;
                if (q.off[x] == q.off[x + 1]) continue;
                const int i = atomicAdd(&counter[x], 1);
                if (q.off[x] + i < q.off[x + 1]) {
                    w = q.off[x] + i;
                    break;
                }
            }
            next = w;
        }
        __syncthreads();
        // the same value in every lane: read it as a scalar so the problem
        // fields live in scalar registers and the tile routines' in-bounds
        // tests are uniform branches
        const int w = __builtin_amdgcn_readfirstlane(next);
        if (w < 0) break;  // every queue is drained
        const int4 t = work[w];
        const GroupedProblemDev p = probs[t.x];
        const int shape = t.w;
        const int tr = p.flags & (GP_TA | GP_TB);
        if constexpr (FORM == GP_FORM_F32) {
            if (tr == 0) f32_tile_shape<false, false>(p, shape, t.y, t.z, gsmem);
            else if (tr == GP_TA) f32_tile_shape<true, false>(p, shape, t.y, t.z, gsmem);
            else if (tr == GP_TB) f32_tile_shape<false, true>(p, shape, t.y, t.z, gsmem);
            else f32_tile_shape<true, true>(p, shape, t.y, t.z, gsmem);
        } else if constexpr (FORM == GP_FORM_BF16) {
            __bf16* hs = (__bf16*)gsmem;
            if (tr == 0) bf16_tile_shape<false, false>(p, shape, t.y, t.z, hs);
            else if (tr == GP_TA) bf16_tile_shape<true, false>(p, shape, t.y, t.z, hs);
            else if (tr == GP_TB) bf16_tile_shape<false, true>(p, shape, t.y, t.z, hs);
            else bf16_tile_shape<true, true>(p, shape, t.y, t.z, hs);
        } else {
            __bf16* hs = (__bf16*)gsmem;
            if (tr == 0) x3_tile_shape<false, false>(p, shape, t.y, t.z, hs);
            else if (tr == GP_TA) x3_tile_shape<true, false>(p, shape, t.y, t.z, hs);
            else if (tr == GP_TB) x3_tile_shape<false, true>(p, shape, t.y, t.z, hs);
            else x3_tile_shape<true, true>(p, shape, t.y, t.z, hs);
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&counter[GP_XCDS], 1) == (int)gridDim.x - 1) {
            for (int x = 0; x <= GP_XCDS; ++x) atomicExch(&counter[x], 0);
        }
    }
}

size_t grouped_smem_bytes(int form) {
    switch (form) {
        case GP_FORM_F32: return grouped_tile_bytes<GP_FORM_F32>() + 16;
        case GP_FORM_BF16: return grouped_tile_bytes<GP_FORM_BF16>() + 16;
        default: return grouped_tile_bytes<GP_FORM_X3>() + 16;
    }
}

int grouped_wgs_per_cu(int form) {
    switch (form) {
        case GP_FORM_F32: return grouped_wgs_per_cu<GP_FORM_F32>();
        case GP_FORM_BF16: return grouped_wgs_per_cu<GP_FORM_BF16>();
        default: return grouped_wgs_per_cu<GP_FORM_X3>();
    }
}

static const void* grouped_kernel_fn(int form) {
    switch (form) {
        case GP_FORM_F32: return (const void*)grouped_gemm_kernel<GP_FORM_F32>;
        case GP_FORM_BF16: return (const void*)grouped_gemm_kernel<GP_FORM_BF16>;
        default: return (const void*)grouped_gemm_kernel<GP_FORM_X3>;
    }
}

int grouped_gemm_prepare(int form) {
    if (form < GP_FORM_F32 || form > GP_FORM_X3) return ST_INVALID;
    const size_t smem = grouped_smem_bytes(form);
    if (smem > 65536) GAT_HIP_TRY(allow_dynamic_lds(grouped_kernel_fn(form), smem));
    return ST_OK;
}

int launch_grouped_gemm(int form, const GroupedProblemDev* probs, const int4* work, const GroupedQueues& q,
                        int* counter, int grid, hipStream_t st) {
    const int rc = grouped_gemm_prepare(form);
    if (rc) return rc;
    const size_t smem = grouped_smem_bytes(form);
    if (form == GP_FORM_F32)
        hipLaunchKernelGGL(grouped_gemm_kernel<GP_FORM_F32>, dim3(grid), dim3(256), smem, st, probs, work, q, counter);
    else if (form == GP_FORM_BF16)
        hipLaunchKernelGGL(grouped_gemm_kernel<GP_FORM_BF16>, dim3(grid), dim3(256), smem, st, probs, work, q,
                           counter);
    else
        hipLaunchKernelGGL(grouped_gemm_kernel<GP_FORM_X3>, dim3(grid), dim3(256), smem, st, probs, work, q, counter);
    GAT_HIP_TRY(hipGetLastError());
    return ST_OK;
}

}  // namespace gat
