#include "checks.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <thread>

namespace gat {
namespace checks {

int hostThreads() {
    unsigned n = std::thread::hardware_concurrency();
    const char* e = std::getenv("OMP_NUM_THREADS");
    if (e && std::atoi(e) > 0) n = (unsigned)std::atoi(e);
    if (n == 0) n = 1;
    if (n > 64) n = 64;
    return (int)n;
}

template <typename F>
static void parallelRows(size_t rows, F fn) {
    const int nt = std::min<int>(hostThreads(), (int)std::max<size_t>(rows, 1));
    if (nt <= 1) {
        fn(0, rows);
        return;
    }
    std::vector<std::thread> th;
    const size_t chunk = (rows + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const size_t a = t * chunk, b = std::min(rows, a + chunk);
        if (a >= b) break;
        th.emplace_back(fn, a, b);
    }
    for (auto& x : th) x.join();
}

template <typename T>
T posrand() {
    T a;
    while (0 == (a = (T)drand48()))
        ;
    return a;
}

template <typename T>
void fillrand(T* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = posrand<T>();
}

template <typename T>
bool checkArray(const T* p, size_t n, T v) {
    bool good = true;
    for (size_t i = 0; i < n; ++i)
        if (v != p[i]) good = false;
    return good;
}

template <typename T>
long double absdiff(const T* a, const T* b, size_t n) {
    long double acc = 0;
    for (size_t i = 0; i < n; ++i) {
        const T d = a[i] - b[i];
        acc += fabsl((long double)d);
    }
    return acc;
}

template <typename T>
bool paranoidPass(const T* a, const T* b, size_t n, long double* diff) {
    const long double d = absdiff(a, b, n);
    if (diff) *diff = d;
    return d < (long double)(1.0 / (double)n);
}

// acc + a*b with the product rounded first (no contraction).
template <typename T>
static inline T mad2(T acc, T a, T b) {
#pragma clang fp contract(off)
    const T p = a * b;
    return acc + p;
}
template <typename T>
static inline T epi2(T alpha, T acc, T beta, T c) {
#pragma clang fp contract(off)
    const T u = alpha * acc;
    const T v = beta * c;
    return u + v;
}

template <typename T>
void refMatmul(size_t packed, bool gemm, bool tA, bool tB, size_t M, size_t N, size_t K, T alpha, T beta,
               const T* A, const T* B, const T* Cold, T* C) {
    for (size_t p = 0; p < packed; ++p) {
        const T* pa = A + p * M * K;
        const T* pb = B + p * K * N;
        const T* po = Cold ? Cold + p * M * N : nullptr;
        T* pc = C + p * M * N;
        parallelRows(M, [&](size_t r0, size_t r1) {
            std::vector<T> acc(N), bcol(N);
            for (size_t i = r0; i < r1; ++i) {
                std::fill(acc.begin(), acc.end(), (T)0);
                for (size_t k = 0; k < K; ++k) {
                    const T a = tA ? pa[k * M + i] : pa[i * K + k];
                    const T* brow;
                    if (!tB) {
                        brow = pb + k * N;
                    } else {
                        for (size_t j = 0; j < N; ++j) bcol[j] = pb[j * K + k];
                        brow = bcol.data();
                    }
                    for (size_t j = 0; j < N; ++j) acc[j] = mad2<T>(acc[j], a, brow[j]);
                }
                for (size_t j = 0; j < N; ++j)
                    pc[i * N + j] = gemm ? epi2<T>(alpha, acc[j], beta, po[i * N + j]) : acc[j];
            }
        });
    }
}

template <typename T>
void refMatvec(size_t packed, bool gemv, bool tA, size_t M, size_t N, T alpha, T beta, const T* A, const T* x,
               const T* yold, T* y) {
    for (size_t p = 0; p < packed; ++p) {
        const T* pa = A + p * M * N;
        const T* px = x + p * N;
        const T* po = yold ? yold + p * M : nullptr;
        T* py = y + p * M;
        parallelRows(M, [&](size_t r0, size_t r1) {
            for (size_t i = r0; i < r1; ++i) {
                T acc = 0;
                for (size_t j = 0; j < N; ++j) acc = mad2<T>(acc, tA ? pa[j * M + i] : pa[i * N + j], px[j]);
                py[i] = gemv ? epi2<T>(alpha, acc, beta, po[i]) : acc;
            }
        });
    }
}

template <typename T>
void refSaxpy(size_t n, T alpha, const T* x, const T* y, T* z) {
    parallelRows(n, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) z[i] = mad2<T>(y[i], alpha, x[i]);
    });
}

void gemm64Rows(bool tA, bool tB, size_t M, size_t N, size_t K, double alpha, const double* A, const double* B,
                double beta, const double* C0, size_t r0, size_t r1, double* C64, double* MAG) {
    parallelRows(r1 - r0, [&](size_t a, size_t b) {
        std::vector<double> acc(N), mag(N);
        for (size_t r = a; r < b; ++r) {
            const size_t i = r0 + r;
            std::fill(acc.begin(), acc.end(), 0.0);
            std::fill(mag.begin(), mag.end(), 0.0);
            for (size_t k = 0; k < K; ++k) {
                const double av = tA ? A[k * M + i] : A[i * K + k];
                const double aa = std::fabs(av);
                for (size_t j = 0; j < N; ++j) {
                    const double bv = tB ? B[j * K + k] : B[k * N + j];
                    acc[j] += av * bv;
                    mag[j] += aa * std::fabs(bv);
                }
            }
            for (size_t j = 0; j < N; ++j) {
                const double c0 = (beta != 0.0 && C0) ? C0[i * N + j] : 0.0;
                C64[r * N + j] = alpha * acc[j] + beta * c0;
                MAG[r * N + j] = std::fabs(alpha) * mag[j] + std::fabs(beta) * std::fabs(c0);
            }
        }
    });
}

double componentwiseError(const double* got, const double* ref, const double* mag, size_t n) {
    double e = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double d = std::fabs(got[i] - ref[i]);
        if (mag[i] > 0.0) e = std::max(e, d / mag[i]);
        else if (d > 0.0) e = INFINITY;
    }
    return e;
}

#define GAT_INST(T)                                                                                          \
    template T posrand<T>();                                                                                 \
    template void fillrand<T>(T*, size_t);                                                                   \
    template bool checkArray<T>(const T*, size_t, T);                                                        \
    template long double absdiff<T>(const T*, const T*, size_t);                                             \
    template bool paranoidPass<T>(const T*, const T*, size_t, long double*);                                 \
    template void refMatmul<T>(size_t, bool, bool, bool, size_t, size_t, size_t, T, T, const T*, const T*,   \
                               const T*, T*);                                                                \
    template void refMatvec<T>(size_t, bool, bool, size_t, size_t, T, T, const T*, const T*, const T*, T*); \
    template void refSaxpy<T>(size_t, T, const T*, const T*, T*);
GAT_INST(float)
GAT_INST(double)
#undef GAT_INST

}  // namespace checks
}  // namespace gat
