// One macro tile of the BF16 matrix-core GEMMs (v_mfma_f32_32x32x16_bf16),
// shared by the per-problem kernel and the persistent grouped kernel.
//
//   MODE_BF16: A and B are BF16 arrays.
//   MODE_X3:   A and B are FP32; while a K tile is staged, every element is
//              split into hi = bf16(a) and lo = bf16(a - hi) (both rounded to
//              nearest even; a - hi is exact in FP32), and the product is
//              accumulated as hi*hi + hi*lo + lo*hi ("split-BF16x3").  The
//              dropped lo*lo term and the split residuals are below
//              2^-16 |a||b| per product.
//
// Decomposition as in the FP32 MFMA kernel: WM x WN wavefronts per group,
// TM x TN 32x32 blocks per wave, K staged BK deep through double-buffered
// LDS; every operand plane is stored k-contiguous ([row][k], row stride
// BK+8 bf16 = a multiple of 16 bytes) so one ds_read_b128 hands a lane the
// eight k values of its MFMA operand: lane l holds row l%32 and k
// kk + 8*(l/32) .. +7 of a 16-deep step.
#pragma once

#include "mfma_common.hpp"

namespace gat {

enum { MODE_BF16 = 0, MODE_X3 = 1 };

__device__ __forceinline__ void split_bf16(float a, __bf16& hi, __bf16& lo) {
    hi = (__bf16)a;
    lo = (__bf16)(a - (float)hi);
}

// split_bf16 of two elements, bit patterns packed first-in-low-half: one
// paired conversion per plane, and the hi values widened back to FP32 by a
// shift (first) and a mask (second) instead of a conversion each.
__device__ __forceinline__ void split_bf16x2(float a0, float a1, unsigned& hi, unsigned& lo) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    typedef __bf16 b2v __attribute__((ext_vector_type(2)));
    hi = __builtin_bit_cast(unsigned, __builtin_convertvector((f2v){a0, a1}, b2v));
    const float h0 = __builtin_bit_cast(float, hi << 16), h1 = __builtin_bit_cast(float, hi & 0xFFFF0000u);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f2v){a0 - h0, a1 - h1}, b2v));
}

// Global -> register -> LDS staging of a ROWS x BK tile for one operand.
//   direct (k-contiguous source): element (r, k) at src[(r0 + r) * ld + k0 + k]
//   transposed (row-contiguous):  element (r, k) at src[(k0 + k) * ld + r0 + r]
// A transposed tile is staged in blocks of KB consecutive k by W consecutive
// rows (KB 16-byte loads, one per k); the store transposes a block in
// registers and writes each of its W rows as one KB-element vector, so LDS
// takes 4-, 8- or 16-byte stores instead of one store per element.  Lanes
// take consecutive k-blocks first: the KB loads of a wavefront read whole
// row segments, and one store instruction spreads over all LDS banks.
// KB is the largest of 8, 4, 2 that divides the per-thread vector count
// and BK (1 -- one store per element -- when none does).
template <int NV, int BK>
struct TransBlock {
    static constexpr int KB = (NV % 8 == 0 && BK % 8 == 0)   ? 8
                              : (NV % 4 == 0 && BK % 4 == 0) ? 4
                              : (NV % 2 == 0 && BK % 2 == 0) ? 2
                                                             : 1;
};

// bf16 bit patterns of two elements packed into one dword (first in the low half)
__device__ __forceinline__ unsigned pack_bf16(unsigned lo, unsigned hi) { return (lo & 0xFFFFu) | (hi << 16); }

// Writes KB bf16 values (bit patterns in the low halves of b[]) to dst
template <int KB>
__device__ __forceinline__ void store_kvec(__bf16* dst, const unsigned (&b)[KB]) {
    if constexpr (KB == 1) {
        *(unsigned short*)dst = (unsigned short)b[0];
    } else if constexpr (KB == 2) {
        *(unsigned*)dst = pack_bf16(b[0], b[1]);
    } else if constexpr (KB == 4) {
        *(u32x2*)dst = u32x2{pack_bf16(b[0], b[1]), pack_bf16(b[2], b[3])};
    } else {
        *(u32x4*)dst = u32x4{pack_bf16(b[0], b[1]), pack_bf16(b[2], b[3]), pack_bf16(b[4], b[5]),
                             pack_bf16(b[6], b[7])};
    }
}

__device__ __forceinline__ unsigned bf16_bits(__bf16 h) { return (unsigned)__builtin_bit_cast(unsigned short, h); }

// Writes NW packed dwords (two bf16 each) to dst as one 4-, 8- or 16-byte store
template <int NW>
__device__ __forceinline__ void store_words(__bf16* dst, const unsigned (&w)[NW]) {
    if constexpr (NW == 1) {
        *(unsigned*)dst = w[0];
    } else if constexpr (NW == 2) {
        *(u32x2*)dst = u32x2{w[0], w[1]};
    } else {
        *(u32x4*)dst = u32x4{w[0], w[1], w[2], w[3]};
    }
}

template <int MODE, int ROWS, int BK, int NT, bool TRANS>
struct StageIO;

// ---- BF16 inputs: 8 bf16 (16 bytes) per vector load
template <int ROWS, int BK, int NT, bool TRANS>
struct StageIO<MODE_BF16, ROWS, BK, NT, TRANS> {
    typedef __bf16 src_t;
    static constexpr int W = 8;
    static constexpr int NV = ROWS * BK / W / NT;
    static_assert(NV >= 1 && NV * W * NT == ROWS * BK, "tile must split into 16-byte vectors per thread");
    static constexpr int LDK = BK + 8;
    static constexpr int KB = TRANS ? TransBlock<NV, BK>::KB : 1;
    u32x4 reg[NV];

    // element (row, k) offset of vector q of this thread in the tile
    __device__ static __forceinline__ void coords(int q, int tid, int& r, int& k) {
        if (!TRANS) {
            const int idx = q * NT + tid;
            r = idx / (BK / W);
            k = (idx % (BK / W)) * W;
        } else {
            const int bi = (q / KB) * NT + tid;  // block index; q % KB is the k within the block
            k = (bi % (BK / KB)) * KB + q % KB;
            r = (bi / (BK / KB)) * W;
        }
    }

    // full: the whole tile is in bounds; otherwise each vector is checked on
    // its own (out-of-range elements read as zero, an in-range vector is one
    // 16-byte load when `vec`).
    __device__ __forceinline__ void load(const __bf16* __restrict__ src, int64_t ld, int r0, int k0, int rlim,
                                         int klim, bool full, bool vec, int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            int r, k;
            coords(q, tid, r, k);
            const int gr = r0 + r, gk = k0 + k;
            const __bf16* p = !TRANS ? src + (int64_t)gr * ld + gk : src + (int64_t)gk * ld + gr;
            const bool vin = !TRANS ? (gr < rlim && gk + W <= klim) : (gk < klim && gr + W <= rlim);
            if (full) {  // uniform: interior tiles
                reg[q] = gload<u32x4>(p);
            } else if (vec && vin) {
                reg[q] = gload<u32x4>(p);
            } else {
                __bf16 t[W];
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    const bool in = !TRANS ? (gr < rlim && gk + e < klim) : (gk < klim && gr + e < rlim);
                    t[e] = in ? gload<__bf16>(p + e) : (__bf16)0.0f;
                }
                reg[q] = *(const u32x4*)t;
            }
        }
    }
    // Interior tiles (every stage whole): each vector's element offset from
    // the stage's first element, computed once per tile; a stage is then one
    // 16-byte load per vector.
    int64_t goff[NV];
    __device__ __forceinline__ void prep(int64_t ld, int r0, int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            int r, k;
            coords(q, tid, r, k);
            goff[q] = !TRANS ? (int64_t)(r0 + r) * ld + k : (int64_t)k * ld + r0 + r;
        }
    }
    __device__ __forceinline__ void load_at(const __bf16* __restrict__ base) {
#pragma unroll
        for (int q = 0; q < NV; ++q) reg[q] = gload<u32x4>(base + goff[q]);
    }
    __device__ __forceinline__ void store(__bf16* hi, __bf16* /*lo*/, int tid) const {
        if constexpr (!TRANS) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                int r, k;
                coords(q, tid, r, k);
                *(u32x4*)(hi + r * LDK + k) = reg[q];
            }
        } else {
#pragma unroll
            for (int qb = 0; qb < NV; qb += KB) {
                int r, k;
                coords(qb, tid, r, k);
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    unsigned b[KB];
#pragma unroll
                    for (int j = 0; j < KB; ++j) b[j] = reg[qb + j][e >> 1] >> (16 * (e & 1));
                    store_kvec<KB>(hi + (r + e) * LDK + k, b);
                }
            }
        }
    }
};

// ---- FP32 inputs split into hi/lo planes: 4 floats per vector load
template <int ROWS, int BK, int NT, bool TRANS>
struct StageIO<MODE_X3, ROWS, BK, NT, TRANS> {
    typedef float src_t;
    static constexpr int W = 4;
    static constexpr int NV = ROWS * BK / W / NT;
    static_assert(NV >= 1 && NV * W * NT == ROWS * BK, "tile must split into 16-byte vectors per thread");
    static constexpr int LDK = BK + 8;
    static constexpr int KB = TRANS ? TransBlock<NV, BK>::KB : 1;
    f4v reg[NV];

    __device__ static __forceinline__ void coords(int q, int tid, int& r, int& k) {
        if (!TRANS) {
            const int idx = q * NT + tid;
            r = idx / (BK / W);
            k = (idx % (BK / W)) * W;
        } else {
            const int bi = (q / KB) * NT + tid;
            k = (bi % (BK / KB)) * KB + q % KB;
            r = (bi / (BK / KB)) * W;
        }
    }

    __device__ __forceinline__ void load(const float* __restrict__ src, int64_t ld, int r0, int k0, int rlim,
                                         int klim, bool full, bool vec, int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            int r, k;
            coords(q, tid, r, k);
            const int gr = r0 + r, gk = k0 + k;
            const float* p = !TRANS ? src + (int64_t)gr * ld + gk : src + (int64_t)gk * ld + gr;
            const bool vin = !TRANS ? (gr < rlim && gk + W <= klim) : (gk < klim && gr + W <= rlim);
            if (full) {  // uniform: interior tiles
                reg[q] = gload<f4v>(p);
            } else if (vec && vin) {
                reg[q] = gload<f4v>(p);
            } else {
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    const bool in = !TRANS ? (gr < rlim && gk + e < klim) : (gk < klim && gr + e < rlim);
                    reg[q][e] = in ? gload<float>(p + e) : 0.0f;
                }
            }
        }
    }
    // interior tiles: as the BF16 form's
    int64_t goff[NV];
    __device__ __forceinline__ void prep(int64_t ld, int r0, int tid) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            int r, k;
            coords(q, tid, r, k);
            goff[q] = !TRANS ? (int64_t)(r0 + r) * ld + k : (int64_t)k * ld + r0 + r;
        }
    }
    __device__ __forceinline__ void load_at(const float* __restrict__ base) {
#pragma unroll
        for (int q = 0; q < NV; ++q) reg[q] = gload<f4v>(base + goff[q]);
    }
    __device__ __forceinline__ void store(__bf16* hi, __bf16* lo, int tid) const {
        if constexpr (!TRANS) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                int r, k;
                coords(q, tid, r, k);
                unsigned h0, l0, h1, l1;
                split_bf16x2(reg[q][0], reg[q][1], h0, l0);
                split_bf16x2(reg[q][2], reg[q][3], h1, l1);
                *(u32x2*)(hi + r * LDK + k) = u32x2{h0, h1};
                *(u32x2*)(lo + r * LDK + k) = u32x2{l0, l1};
            }
        } else {
#pragma unroll
            for (int qb = 0; qb < NV; qb += KB) {
                int r, k;
                coords(qb, tid, r, k);
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    if constexpr (KB == 1) {
                        __bf16 h, l;
                        split_bf16(reg[qb][e], h, l);
                        *(unsigned short*)(hi + (r + e) * LDK + k) = (unsigned short)bf16_bits(h);
                        *(unsigned short*)(lo + (r + e) * LDK + k) = (unsigned short)bf16_bits(l);
                    } else {
                        // k pairs (j, j + 1) split together, packed first-in-low-half
                        unsigned wh[KB / 2], wl[KB / 2];
#pragma unroll
                        for (int j = 0; j < KB; j += 2)
                            split_bf16x2(reg[qb + j][e], reg[qb + j + 1][e], wh[j / 2], wl[j / 2]);
                        store_words<KB / 2>(hi + (r + e) * LDK + k, wh);
                        store_words<KB / 2>(lo + (r + e) * LDK + k, wl);
                    }
                }
            }
        }
    }
};

template <int WM, int WN, int TM, int TN, int BK, int MODE>
struct Bf16Shape {
    static constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    static constexpr size_t SMEM = 2 * (size_t)(MODE == MODE_X3 ? 2 : 1) * (BM + BN) * (BK + 8) * sizeof(__bf16);
};

// Computes the BM x BN tile at (m0, n0); called by all WM*WN*64 threads of
// the group, ends after a barrier (LDS free for the next tile).
template <int WM, int WN, int TM, int TN, int BK, int MODE, bool TA, bool TB>
__device__ __forceinline__ void bf16_tile(const void* __restrict__ Av, const void* __restrict__ Bv,
                                          float* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb,
                                          int64_t ldc, float alpha, float beta, int m0, int n0, bool vec_ok,
                                          __bf16* smem) {
    constexpr int NT = WM * WN * 64;
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    constexpr int LDK = BK + 8;
    constexpr int NPL = MODE == MODE_X3 ? 2 : 1;  // operand planes (hi[, lo])
    static_assert(BK % 16 == 0, "BK must cover whole 16-deep MFMA steps");
    typedef StageIO<MODE, BM, BK, NT, TA> IoA;
    typedef StageIO<MODE, BN, BK, NT, !TB> IoB;
    typedef typename IoA::src_t src_t;
    constexpr int PLANE_A = BM * LDK, PLANE_B = BN * LDK;
    constexpr int BUF = NPL * (PLANE_A + PLANE_B);
    const src_t* A = (const src_t*)Av;
    const src_t* B = (const src_t*)Bv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const bool full_mn = vec_ok && (m0 + BM <= M) && (n0 + BN <= N);

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    IoA ioa;
    IoB iob;
    const int KT = (K + BK - 1) / BK;
    // interior tile with whole stages: per-vector offsets computed once
    const bool interior = full_mn && K % BK == 0;
    if (interior) {
        ioa.prep(lda, m0, tid);
        iob.prep(ldb, n0, tid);
    }
    auto fetch = [&](int t) {
        const int k0 = t * BK;
        if (interior) {
            // op(A) rows are k-contiguous unless A is transposed; op(B)^T rows
            // are k-contiguous when B is transposed
            ioa.load_at(A + (TA ? (int64_t)k0 * lda : (int64_t)k0));
            iob.load_at(B + (TB ? (int64_t)k0 : (int64_t)k0 * ldb));
            return;
        }
        const bool full = full_mn && (k0 + BK <= K);
        ioa.load(A, lda, m0, k0, M, K, full, vec_ok, tid);
        iob.load(B, ldb, n0, k0, N, K, full, vec_ok, tid);
    };
    auto sstore = [&](__bf16* buf) {
        ioa.store(buf, buf + PLANE_A + PLANE_B, tid);
        iob.store(buf + PLANE_A, buf + 2 * PLANE_A + PLANE_B, tid);
    };

    fetch(0);
    sstore(smem);
    __syncthreads();

    const int arow = wm * TM * 32 + (lane & 31);
    const int brow = wn * TN * 32 + (lane & 31);
    const int kq = 8 * (lane >> 5);

    for (int t = 0; t < KT; ++t) {
        const __bf16* Ah = smem + (t & 1) * BUF;
        const __bf16* Bh = Ah + PLANE_A;
        const __bf16* Al = Ah + PLANE_A + PLANE_B;
        const __bf16* Bl = Al + PLANE_A;
        if (t + 1 < KT) fetch(t + 1);
        // interleave the LDS fragment reads with the matrix instructions
        __builtin_amdgcn_iglp_opt(0);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 ah[TM], bh[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) ah[i] = *(const bf16x8*)(Ah + (arow + i * 32) * LDK + kk + kq);
#pragma unroll
            for (int j = 0; j < TN; ++j) bh[j] = *(const bf16x8*)(Bh + (brow + j * 32) * LDK + kk + kq);
            if (MODE == MODE_X3) {
                bf16x8 al[TM], bl[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) al[i] = *(const bf16x8*)(Al + (arow + i * 32) * LDK + kk + kq);
#pragma unroll
                for (int j = 0; j < TN; ++j) bl[j] = *(const bf16x8*)(Bl + (brow + j * 32) * LDK + kk + kq);
                // small terms first, then the leading product
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
        if (t + 1 < KT) sstore(smem + ((t + 1) & 1) * BUF);
        __syncthreads();
    }

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

}  // namespace gat
