// The reference's own result checks and host products (OCLAppUtil.hpp:28-61,
// 138-237; KernelMatmulBuffer.hpp:108-147; KernelMatvecBuffer.hpp:107-142;
// KernelSaxpyBuffer.hpp:115-133), plus an FP64 host product for kernels whose
// summation order differs from the reference's.
#pragma once

#include <cstddef>
#include <vector>

namespace gat {
namespace checks {

// posrand()/fillrand(): libc drand48, the stream the reference draws from.
template <typename T>
T posrand();
template <typename T>
void fillrand(T* out, size_t n);

// exact test-pattern compare (checkArray)
template <typename T>
bool checkArray(const T* p, size_t n, T value);

// paranoid rule: sum |a-b| (difference in T, sum in long double) < 1/n
template <typename T>
long double absdiff(const T* a, const T* b, size_t n);
template <typename T>
bool paranoidPass(const T* a, const T* b, size_t n, long double* diff);

// Reference host products, bit-for-bit: every output is a k- (or j-)
// ascending chain of separately rounded multiply and add in T, then the
// GEMM/GEMV epilogue alpha*acc + beta*old.  Multi-threaded over rows.
template <typename T>
void refMatmul(size_t packed, bool gemm, bool tA, bool tB, size_t M, size_t N, size_t K, T alpha, T beta,
               const T* A, const T* B, const T* Cold, T* C);
template <typename T>
void refMatvec(size_t packed, bool gemv, bool tA, size_t M, size_t N, T alpha, T beta, const T* A, const T* x,
               const T* yold, T* y);
template <typename T>
void refSaxpy(size_t n, T alpha, const T* x, const T* y, T* z);

// FP64 product of rows [r0, r1) (all columns) of alpha*op(A)op(B)+beta*C0 and
// the magnitude |alpha||op(A)||op(B)| + |beta||C0|; inputs given as double.
void gemm64Rows(bool tA, bool tB, size_t M, size_t N, size_t K, double alpha, const double* A,
                const double* B, double beta, const double* C0, size_t r0, size_t r1, double* C64, double* MAG);
// max componentwise error |got - ref| / mag over n elements
double componentwiseError(const double* got, const double* ref, const double* mag, size_t n);

int hostThreads();

}  // namespace checks
}  // namespace gat
