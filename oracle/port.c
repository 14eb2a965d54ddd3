/*
 * port.c -- CPU port of the reference's buffer-kernel arithmetic, timed as
 * the CPU baseline.  TEST INFRASTRUCTURE: only bench.py's CPU-baseline leg
 * and the tests call it; the library never links it.
 *
 * Every output element is computed with the reference's own arithmetic:
 * a k- (j-) ascending chain of separately rounded multiply and add
 * (KernelMatmulBuffer.hpp:259-337, KernelMatvecBuffer.hpp:280-320,
 * KernelSaxpyBuffer.hpp:169-220 with the two roundings of its host check).
 * Built -O3 with -ffp-contract=off, so vectorisation runs across independent
 * output elements without fusing or reordering any element's operations;
 * each function is cloned for AVX-512, AVX2 and baseline x86-64 and the
 * loader picks the widest the host supports (the GPU box's CPU is not known
 * when this is built).  Threads: OpenMP over output rows.
 */
#include <stddef.h>
#include <stdlib.h>
#include <omp.h>

#define PORT_CLONES __attribute__((target_clones("avx512f", "avx2", "default")))

/* C rows [row0, row1) of C = alpha*op(A)op(B) + beta*C (gemm) or
 * C = op(A)op(B): every element is a k-ascending chain of unfused
 * multiply-then-add in float, the arithmetic of KernelMatmulBuffer.hpp:259-337
 * with its host check's rounding.  Loops run i, k, j so that the j loop
 * vectorises across independent output elements without changing any
 * element's k order.  Returns nothing; the caller times it. */
/* op(B) = B^T (B stored N x K): transposed once into a K x N copy, in
 * 64 x 64 blocks, so the product's j loop reads rows; this layout change is
 * part of the timed work (O(NK) next to O(MNK)). */
PORT_CLONES
static void port_transpose(const float *B, float *Bt, size_t N, size_t K)
{
    _Pragma("omp parallel for schedule(static) collapse(2)")
    for (size_t j0 = 0; j0 < N; j0 += 64)
        for (size_t k0 = 0; k0 < K; k0 += 64) {
            const size_t j1 = (j0 + 64 < N) ? j0 + 64 : N, k1 = (k0 + 64 < K) ? k0 + 64 : K;
            for (size_t j = j0; j < j1; j++)
                for (size_t k = k0; k < k1; k++) Bt[k * N + j] = B[j * K + k];
        }
}

PORT_CLONES
void ora_port_matmul_f32_rows(int gemm, int transA, int transB,
                              size_t M, size_t N, size_t K,
                              float alpha, float beta,
                              const float *A, const float *B, float *C,
                              size_t row0, size_t row1)
{
    const size_t JB = 512;
    float *Bt = NULL;
    if (transB) {
        Bt = (float *)malloc(K * N * sizeof(float));
        if (!Bt) return;
        port_transpose(B, Bt, N, K);
        B = Bt;
    }
    _Pragma("omp parallel")
    {
        float acc[512];
        _Pragma("omp for schedule(dynamic, 1) collapse(2)")
        for (size_t i = row0; i < row1; i++)
            for (size_t j0 = 0; j0 < N; j0 += JB) {
                const size_t jn = (j0 + JB < N) ? JB : N - j0;
                for (size_t j = 0; j < jn; j++) acc[j] = 0.0f;
                for (size_t k = 0; k < K; k++) {
                    const float a = transA ? A[k * M + i] : A[i * K + k];
                    const float *brow = B + k * N + j0;
                    for (size_t j = 0; j < jn; j++) {
                        const float prod = a * brow[j];
                        acc[j] = acc[j] + prod;
                    }
                }
                float *crow = C + i * N + j0;
                if (gemm) {
                    for (size_t j = 0; j < jn; j++) {
                        const float x = alpha * acc[j];
                        const float y = beta * crow[j];
                        crow[j] = x + y;
                    }
                } else {
                    for (size_t j = 0; j < jn; j++) crow[j] = acc[j];
                }
            }
    }
    free(Bt);
}

/* y = alpha*A*x + beta*y (gemv) over rows [row0,row1), j-ascending unfused. */
PORT_CLONES
void ora_port_matvec_f32_rows(int gemv, size_t N, float alpha, float beta,
                              const float *A, const float *x, float *y,
                              size_t row0, size_t row1)
{
    _Pragma("omp parallel for schedule(static)")
    for (size_t i = row0; i < row1; i++) {
        float acc = 0.0f;
        const float *arow = A + i * N;
        for (size_t j = 0; j < N; j++) {
            const float prod = arow[j] * x[j];
            acc = acc + prod;
        }
        if (gemv) {
            const float u = alpha * acc;
            const float v = beta * y[i];
            y[i] = u + v;
        } else {
            y[i] = acc;
        }
    }
}

/* z = alpha*x + y over [i0, i1): product rounded, then the add rounded. */
PORT_CLONES
void ora_port_saxpy_f32(size_t i0, size_t i1, float alpha, const float *x, const float *y, float *z)
{
    _Pragma("omp parallel for schedule(static)")
    for (size_t i = i0; i < i1; i++) {
        const float p = alpha * x[i];
        z[i] = p + y[i];
    }
}
