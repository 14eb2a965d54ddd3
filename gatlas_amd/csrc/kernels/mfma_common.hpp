// Shared device helpers of the matrix-core GEMM kernels.
#pragma once

#include "../kernels.hpp"
#include "gmem.hpp"

namespace gat {

// s_waitcnt operand for gfx9-family targets: vmcnt in bits 3:0 and 15:14,
// expcnt (7: no wait) in 6:4, lgkmcnt in 11:8.  Issued as a builtin, the
// wait is known to the compiler, which then adds none of its own after it.
constexpr int waitcnt_imm(int vm, int lgkm) { return (vm & 15) | ((vm >> 4) << 14) | (7 << 4) | (lgkm << 8); }

typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Work-group id -> (tile_m, tile_n): ids are first regrouped so that each
// of the 8 XCDs (which receive consecutive ids round-robin) owns a
// contiguous range, then walked in column-major supertiles `group_m` tiles
// tall so that concurrently resident tiles share A and B panels in L2.
__device__ __forceinline__ void mfma_tile_coords(int wg, int nwg, int tiles_m, int tiles_n, int group_m, int& tm,
                                                 int& tn) {
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

// 32x32 MFMA output element (lane, accumulator r) -> (row, col) in the block
__device__ __forceinline__ int mfma32_row(int lane, int r) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }
__device__ __forceinline__ int mfma32_col(int lane) { return lane & 31; }

}  // namespace gat
