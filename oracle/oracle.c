/*
 * oracle.c -- CPU oracle for the GEMM / GEMV / SAXPY hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; only tests/, the smoke check and the bench's CPU-baseline leg
 * load it (through oracle/oracle.py), and only as the checker.
 *
 * It restates, from the reference sources, three things:
 *
 *  1. The reference's own result checks and their input stream
 *     (OCLAppUtil.hpp:28-61, 212-237): the 48-bit drand48 LCG that feeds
 *     posrand(), the paranoid CPU products that the reference computes in
 *     KernelMatmulBuffer.hpp:108-147, KernelMatvecBuffer.hpp:107-142 and
 *     KernelSaxpyBuffer.hpp:115-133, and the absdiff / checkArray rules.
 *     These are restated bit-for-bit: scalar precision, the reference's loop
 *     order, and separate multiply and add roundings (this file must be
 *     compiled with -ffp-contract=off, which oracle/Makefile does).
 *
 *  2. An FP64 host oracle for every product (general form
 *     C = alpha*op(A)*op(B) + beta*C, y = alpha*op(A)*x + beta*y,
 *     z = alpha*x + y), blocked and OpenMP-parallel, with the error metrics
 *     used to judge the fast paths: componentwise
 *     max |C - C64| / (|alpha| |op(A)| |op(B)| + |beta| |C0|) and normwise
 *     ||C - C64||_F / ||C64||_F.  A row-sampled variant serves large sizes.
 *
 *  3. A multi-threaded CPU port of the reference buffer-kernel arithmetic
 *     (k-ascending accumulation per output element, unfused multiply/add,
 *     epilogue alpha*acc + beta*C) used as the timed CPU baseline.
 *
 * Pinning (tests/test_oracle.py): glibc's drand48 for the LCG stream;
 * integer-valued inputs where every float operation is exact; hand-built
 * inputs whose float result depends on the summation order and on fused vs
 * unfused rounding; numpy float64 products; transposition invariance.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* drand48 stream (OCLAppUtil.hpp:28-39 posrand/fillrand)              */
/* X_{n+1} = (0x5DEECE66D * X_n + 0xB) mod 2^48, value = X_{n+1} / 2^48 */
/* glibc starts an unseeded stream from X = 0.                         */
/* ------------------------------------------------------------------ */

#define LCG_A 0x5DEECE66DULL
#define LCG_C 0xBULL
#define LCG_MASK ((1ULL << 48) - 1)

double ora_lcg48_next(uint64_t *state)
{
    *state = (LCG_A * *state + LCG_C) & LCG_MASK;
    return ldexp((double)*state, -48);
}

/* posrand<float>(): first draw whose float conversion is nonzero. */
float ora_posrand_f32(uint64_t *state)
{
    float a;
    while (0 == (a = (float)ora_lcg48_next(state)))
        ;
    return a;
}

double ora_posrand_f64(uint64_t *state)
{
    double a;
    while (0 == (a = ora_lcg48_next(state)))
        ;
    return a;
}

void ora_fillrand_f32(uint64_t *state, float *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = ora_posrand_f32(state);
}

void ora_fillrand_f64(uint64_t *state, double *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = ora_posrand_f64(state);
}

/* ------------------------------------------------------------------ */
/* Reference paranoid products, bit-for-bit.                          */
/* ------------------------------------------------------------------ */

/* KernelMatmulBuffer.hpp:117-143.  C is packed x M x N row-major.  A is
 * M x K row-major, or K x M row-major when transA (column-major A).  B is
 * K x N row-major, or N x K row-major when transB.  Accumulation is
 * C[i][j] += a*b for k ascending, in scalar precision; GEMM form then sets
 * C = alpha*C + beta*Cold. */
#define DEF_REF_MATMUL(NAME, T)                                                  \
void NAME(int packed, int gemm, int transA, int transB,                          \
          size_t M, size_t N, size_t K, T alpha, T beta,                         \
          const T *A, const T *B, const T *Cold, T *C)                           \
{                                                                                \
    for (int p = 0; p < packed; p++) {                                           \
        const T *pa = A + (size_t)p * M * K;                                     \
        const T *pb = B + (size_t)p * K * N;                                     \
        T *pc = C + (size_t)p * M * N;                                           \
        _Pragma("omp parallel for schedule(static)")                             \
        for (size_t i = 0; i < M; i++)                                           \
            for (size_t j = 0; j < N; j++) {                                     \
                T acc = 0;                                                       \
                for (size_t k = 0; k < K; k++) {                                 \
                    const T a = transA ? pa[k * M + i] : pa[i * K + k];          \
                    const T b = transB ? pb[j * K + k] : pb[k * N + j];          \
                    const T prod = a * b;                                        \
                    acc = acc + prod;                                            \
                }                                                                \
                pc[i * N + j] = acc;                                             \
            }                                                                    \
        if (gemm) {                                                              \
            const T *po = Cold + (size_t)p * M * N;                              \
            for (size_t i = 0; i < M * N; i++) {                                 \
                const T x = alpha * pc[i];                                       \
                const T y = beta * po[i];                                        \
                pc[i] = x + y;                                                   \
            }                                                                    \
        }                                                                        \
    }                                                                            \
}
DEF_REF_MATMUL(ora_ref_matmul_f32, float)
DEF_REF_MATMUL(ora_ref_matmul_f64, double)

/* KernelMatvecBuffer.hpp:118-133.  y[i] += a(i,j) * x[j], j ascending;
 * A is M x N row-major, or N x M row-major when transA. */
#define DEF_REF_MATVEC(NAME, T)                                                  \
void NAME(int packed, int gemv, int transA, size_t M, size_t N, T alpha, T beta, \
          const T *A, const T *x, const T *yold, T *y)                           \
{                                                                                \
    for (int p = 0; p < packed; p++) {                                           \
        const T *pa = A + (size_t)p * M * N;                                     \
        const T *px = x + (size_t)p * N;                                         \
        T *py = y + (size_t)p * M;                                               \
        _Pragma("omp parallel for schedule(static)")                             \
        for (size_t i = 0; i < M; i++) {                                         \
            T acc = 0;                                                           \
            for (size_t j = 0; j < N; j++) {                                     \
                const T a = transA ? pa[j * M + i] : pa[i * N + j];              \
                const T prod = a * px[j];                                        \
                acc = acc + prod;                                                \
            }                                                                    \
            py[i] = acc;                                                         \
        }                                                                        \
        if (gemv) {                                                              \
            const T *po = yold + (size_t)p * M;                                  \
            for (size_t i = 0; i < M; i++) {                                     \
                const T u = alpha * py[i];                                       \
                const T v = beta * po[i];                                        \
                py[i] = u + v;                                                   \
            }                                                                    \
        }                                                                        \
    }                                                                            \
}
DEF_REF_MATVEC(ora_ref_matvec_f32, float)
DEF_REF_MATVEC(ora_ref_matvec_f64, double)

/* KernelSaxpyBuffer.hpp:125-127: z = alpha*x + y. */
void ora_ref_saxpy_f32(size_t n, float alpha, const float *x, const float *y, float *z)
{
    _Pragma("omp parallel for schedule(static)")
    for (size_t i = 0; i < n; i++) {
        const float ax = alpha * x[i];
        z[i] = ax + y[i];
    }
}

void ora_ref_saxpy_f64(size_t n, double alpha, const double *x, const double *y, double *z)
{
    _Pragma("omp parallel for schedule(static)")
    for (size_t i = 0; i < n; i++) {
        const double ax = alpha * x[i];
        z[i] = ax + y[i];
    }
}

/* OCLAppUtil.hpp:47-53: sum of |m1 - m2| where the difference is taken in
 * scalar precision and summed in long double. */
long double ora_absdiff_f32(const float *a, const float *b, size_t n)
{
    long double acc = 0;
    for (size_t i = 0; i < n; i++) {
        const float d = a[i] - b[i];
        acc += fabsl((long double)d);
    }
    return acc;
}

long double ora_absdiff_f64(const double *a, const double *b, size_t n)
{
    long double acc = 0;
    for (size_t i = 0; i < n; i++) {
        const double d = a[i] - b[i];
        acc += fabsl((long double)d);
    }
    return acc;
}

/* OCLAppUtil.hpp:212-237: paranoid pass rule absdiff < 1/length. */
int ora_paranoid_pass_f32(const float *a, const float *b, size_t n)
{
    return ora_absdiff_f32(a, b, n) < (1.0 / (double)n);
}

int ora_paranoid_pass_f64(const double *a, const double *b, size_t n)
{
    return ora_absdiff_f64(a, b, n) < (1.0 / (double)n);
}

/* OCLAppUtil.hpp:55-61: exact test-pattern compare. */
int ora_check_array_f32(const float *p, size_t n, float v)
{
    int good = 1;
    for (size_t i = 0; i < n; i++)
        if (v != p[i]) good = 0;
    return good;
}

int ora_check_array_f64(const double *p, size_t n, double v)
{
    int good = 1;
    for (size_t i = 0; i < n; i++)
        if (v != p[i]) good = 0;
    return good;
}

/* ------------------------------------------------------------------ */
/* FP64 host oracle                                                    */
/* ------------------------------------------------------------------ */

/* op(A)[i][k] = transA ? A[k*lda + i] : A[i*lda + k]  (row-major storage)
 * op(B)[k][j] = transB ? B[j*ldb + k] : B[k*ldb + j]
 * For each sampled row rows[r] (all rows when rows == NULL, nrows = M) and
 * every column j, writes
 *   C64[r*N + j]  = alpha * sum_k op(A)op(B) + beta * C0[row][j]
 *   MAG[r*N + j]  = |alpha| * sum_k |op(A)||op(B)| + |beta| |C0[row][j]|
 * (MAG may be NULL).  Inputs are doubles (float inputs convert exactly).
 * Accumulation is FP64 with k blocked for cache reuse. */
void ora_gemm_f64(int transA, int transB, size_t M, size_t N, size_t K,
                  double alpha, const double *A, size_t lda,
                  const double *B, size_t ldb,
                  double beta, const double *C0, size_t ldc,
                  const long *rows, size_t nrows,
                  double *C64, double *MAG)
{
    const size_t KB = 256;
    if (!rows) nrows = M;
    _Pragma("omp parallel")
    {
        _Pragma("omp for schedule(dynamic, 1)")
        for (size_t r = 0; r < nrows; r++) {
            const size_t i = rows ? (size_t)rows[r] : r;
            double *acc = C64 + r * N;
            double *mag = MAG ? MAG + r * N : NULL;
            for (size_t j = 0; j < N; j++) { acc[j] = 0.0; if (mag) mag[j] = 0.0; }
            for (size_t k0 = 0; k0 < K; k0 += KB) {
                const size_t k1 = (k0 + KB < K) ? k0 + KB : K;
                for (size_t k = k0; k < k1; k++) {
                    const double a = transA ? A[k * lda + i] : A[i * lda + k];
                    if (!transB) {
                        const double *brow = B + k * ldb;
                        for (size_t j = 0; j < N; j++) acc[j] += a * brow[j];
                        if (mag) {
                            const double aa = fabs(a);
                            for (size_t j = 0; j < N; j++) mag[j] += aa * fabs(brow[j]);
                        }
                    } else {
                        for (size_t j = 0; j < N; j++) acc[j] += a * B[j * ldb + k];
                        if (mag) {
                            const double aa = fabs(a);
                            for (size_t j = 0; j < N; j++) mag[j] += aa * fabs(B[j * ldb + k]);
                        }
                    }
                }
            }
            for (size_t j = 0; j < N; j++) {
                const double c0 = (beta != 0.0 && C0) ? C0[i * ldc + j] : 0.0;
                acc[j] = alpha * acc[j] + beta * c0;
                if (mag) mag[j] = fabs(alpha) * mag[j] + fabs(beta) * fabs(c0);
            }
        }
    }
}

/* Same for float inputs: converts on the fly (exact), accumulates in FP64. */
void ora_gemm_f64_from_f32(int transA, int transB, size_t M, size_t N, size_t K,
                           double alpha, const float *A, size_t lda,
                           const float *B, size_t ldb,
                           double beta, const float *C0, size_t ldc,
                           const long *rows, size_t nrows,
                           double *C64, double *MAG)
{
    if (!rows) nrows = M;
    _Pragma("omp parallel")
    {
        _Pragma("omp for schedule(dynamic, 1)")
        for (size_t r = 0; r < nrows; r++) {
            const size_t i = rows ? (size_t)rows[r] : r;
            double *acc = C64 + r * N;
            double *mag = MAG ? MAG + r * N : NULL;
            for (size_t j = 0; j < N; j++) { acc[j] = 0.0; if (mag) mag[j] = 0.0; }
            for (size_t k = 0; k < K; k++) {
                const double a = transA ? (double)A[k * lda + i] : (double)A[i * lda + k];
                const double aa = fabs(a);
                if (!transB) {
                    const float *brow = B + k * ldb;
                    for (size_t j = 0; j < N; j++) acc[j] += a * (double)brow[j];
                    if (mag) for (size_t j = 0; j < N; j++) mag[j] += aa * fabs((double)brow[j]);
                } else {
                    for (size_t j = 0; j < N; j++) acc[j] += a * (double)B[j * ldb + k];
                    if (mag) for (size_t j = 0; j < N; j++) mag[j] += aa * fabs((double)B[j * ldb + k]);
                }
            }
            for (size_t j = 0; j < N; j++) {
                const double c0 = (beta != 0.0 && C0) ? (double)C0[i * ldc + j] : 0.0;
                acc[j] = alpha * acc[j] + beta * c0;
                if (mag) mag[j] = fabs(alpha) * mag[j] + fabs(beta) * fabs(c0);
            }
        }
    }
}

/* y64[i] = alpha * sum_j op(A)[i][j] x[j] + beta * y0[i], op(A) = A (M x N
 * row-major, lda) or A^T with A stored N x M row-major when transA.
 * MAG[i] = |alpha| sum |a||x| + |beta||y0|.  Float inputs, FP64 accumulate. */
void ora_gemv_f64_from_f32(int transA, size_t M, size_t N, double alpha,
                           const float *A, size_t lda, const float *x,
                           double beta, const float *y0, double *y64, double *MAG)
{
    if (!transA) {
        _Pragma("omp parallel for schedule(static)")
        for (size_t i = 0; i < M; i++) {
            double acc = 0.0, mag = 0.0;
            const float *arow = A + i * lda;
            for (size_t j = 0; j < N; j++) {
                acc += (double)arow[j] * (double)x[j];
                mag += fabs((double)arow[j]) * fabs((double)x[j]);
            }
            const double c0 = (beta != 0.0 && y0) ? (double)y0[i] : 0.0;
            y64[i] = alpha * acc + beta * c0;
            if (MAG) MAG[i] = fabs(alpha) * mag + fabs(beta) * fabs(c0);
        }
    } else {
        for (size_t i = 0; i < M; i++) { y64[i] = 0.0; if (MAG) MAG[i] = 0.0; }
        for (size_t j = 0; j < N; j++) {
            const float *acol = A + j * lda;
            const double xj = (double)x[j];
            _Pragma("omp parallel for schedule(static)")
            for (size_t i = 0; i < M; i++) {
                y64[i] += (double)acol[i] * xj;
                if (MAG) MAG[i] += fabs((double)acol[i]) * fabs(xj);
            }
        }
        for (size_t i = 0; i < M; i++) {
            const double c0 = (beta != 0.0 && y0) ? (double)y0[i] : 0.0;
            y64[i] = alpha * y64[i] + beta * c0;
            if (MAG) MAG[i] = fabs(alpha) * MAG[i] + fabs(beta) * fabs(c0);
        }
    }
}

/* Error metrics of an approximate result `got` (n elements, stride 1)
 * against the FP64 oracle `ref` with magnitudes `mag`:
 *   out[0] = max_i |got - ref| / mag  (componentwise, 0/0 counts as 0)
 *   out[1] = ||got - ref||_2 / ||ref||_2  (normwise)
 *   out[2] = max_i |got - ref|  (absolute) */
void ora_error_metrics_f32(const float *got, const double *ref, const double *mag,
                           size_t n, double *out)
{
    double cw = 0.0, num = 0.0, den = 0.0, amax = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double d = fabs((double)got[i] - ref[i]);
        if (d > amax) amax = d;
        if (mag && mag[i] > 0.0) { const double r = d / mag[i]; if (r > cw) cw = r; }
        else if (mag && d > 0.0) cw = INFINITY;
        num += d * d;
        den += ref[i] * ref[i];
    }
    out[0] = cw;
    out[1] = den > 0.0 ? sqrt(num / den) : sqrt(num);
    out[2] = amax;
}

void ora_error_metrics_f64(const double *got, const double *ref, const double *mag,
                           size_t n, double *out)
{
    double cw = 0.0, num = 0.0, den = 0.0, amax = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double d = fabs(got[i] - ref[i]);
        if (d > amax) amax = d;
        if (mag && mag[i] > 0.0) { const double r = d / mag[i]; if (r > cw) cw = r; }
        else if (mag && d > 0.0) cw = INFINITY;
        num += d * d;
        den += ref[i] * ref[i];
    }
    out[0] = cw;
    out[1] = den > 0.0 ? sqrt(num / den) : sqrt(num);
    out[2] = amax;
}

int ora_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
