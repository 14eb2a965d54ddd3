// Row-major GEMV wave routine shared by the single-problem kernel
// (gemv.hip) and the grouped kernel (gemv_grouped.hip), so that one
// problem computes the same bits in either launch.
//
// A wavefront owns ROWS consecutive rows starting at r0 and walks them
// together: every lane loads one 16-byte vector per row per step (1 KiB
// contiguous per row per wave-load, non-temporal since A is streamed
// once), x is read once per step and reused across the ROWS rows, the
// per-lane partial sums are combined by a butterfly reduction and lanes
// 0..ROWS-1 apply the alpha/beta epilogue.
#pragma once

#include "../kernels.hpp"
#include "gmem.hpp"

namespace gat {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T, typename V, int ROWS, int UNROLL>
__device__ __forceinline__ void gemv_rows_wave(int64_t M, int64_t N, T alpha, const T* __restrict__ A,
                                               int64_t lda, const T* __restrict__ x, T beta, T* __restrict__ y,
                                               int vec_ok, int64_t r0, int lane) {
    constexpr int W = sizeof(V) / sizeof(T);
    T acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = 0;
    int64_t c = 0;
    if (vec_ok && r0 + ROWS <= M) {
        const int64_t step = 64 * W;
        const int64_t nfull = (N / (step * UNROLL)) * (step * UNROLL);
        for (; c < nfull; c += step * UNROLL) {
            V av[UNROLL][ROWS], xv[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const int64_t cc = c + u * step + lane * W;
                xv[u] = gload<V>(x + cc);
#pragma unroll
                for (int r = 0; r < ROWS; ++r)
                    av[u][r] = gload_nt<V>(A + (r0 + r) * lda + cc);
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
#pragma unroll
                for (int r = 0; r < ROWS; ++r)
#pragma unroll
                    for (int e = 0; e < W; ++e) acc[r] += av[u][r][e] * xv[u][e];
        }
    }
    // remainder (and unaligned / partial-row-block cases): scalar columns
    for (int64_t cc = c + lane; cc < N; cc += 64) {
        const T xv = gload<T>(x + cc);
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            if (r0 + r < M) acc[r] += gload<T>(A + (r0 + r) * lda + cc) * xv;
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = wave_sum<T>(acc[r]);
    if (lane < ROWS && r0 + lane < M) {
        T s = acc[0];
#pragma unroll
        for (int r = 1; r < ROWS; ++r)
            if (lane == r) s = acc[r];
        T v = alpha * s;
        if (beta != (T)0) v += beta * gload<T>(y + r0 + lane);
        gstore(y + r0 + lane, v);
    }
}

}  // namespace gat
