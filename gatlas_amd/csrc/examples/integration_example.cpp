// The C ABI as a C++ maintainer binds it in place of the reference's
// benchmark plumbing: plain g++ against include/gatlas.h, no HIP headers.
//
//   1. one trial of the reference's buffer SGEMM through the parameter
//      interface (KernelMatmulBuffer<float,4>, GatlasBenchmark.cpp:178-289):
//      test pattern, 10 chained launches, the exact C == K check;
//   2. the direct entry point: C = 1.5*A*B - 0.5*C at 256 x 192 x 320 on
//      device-generated inputs, the result written to the file argv[1]
//      (raw FP32, row-major) for the caller to check.
// Prints "check <0|1> usec <t>" and "gemm <status>"; exit status 0 on success.
#include <cstdio>
#include <vector>

#include "gatlas.h"

int main(int argc, char** argv) {
    gat_ctx* ctx = gat_create(0);
    if (!ctx) {
        std::fprintf(stderr, "gat_create: %s\n", gat_last_error());
        return 1;
    }
    int status = 0;

    // 1. reference parameter vector: [packed, GEMM, M, N, K, transA, transB,
    //    groupHeight, groupWidth, blockHeight, blockWidth, extraParam]
    gat_kernel* k = gat_kernel_new(ctx, GAT_MATMUL, GAT_F32, 4, GAT_BUFFER);
    size_t args[12] = {1, 0, 1024, 1024, 1024, 0, 0, 16, 16, 8, 4, 5};
    if (!k || gat_kernel_set_params(k, args, 12) != GAT_OK || !gat_kernel_valid(k)) {
        std::fprintf(stderr, "kernel: %s\n", gat_last_error());
        status = 1;
    } else {
        double usec = 0, err = 0;
        int ok = 0;
        if (gat_kernel_set_args(k, 1) == GAT_OK && gat_kernel_run(k, 10, &usec) == GAT_OK &&
            gat_kernel_sync_output(k) == GAT_OK)
            ok = gat_kernel_check(k, 0, &err);  // test pattern: C == K exactly
        std::printf("%s check %d usec %.1f\n", gat_kernel_name(k), ok, usec);
        if (ok != 1) status = 1;
    }
    gat_kernel_free(k);

    // 2. direct GEMM on device-resident inputs
    const size_t M = 256, N = 192, K = 320;
    void *A = nullptr, *B = nullptr, *C = nullptr;
    int rc = gat_malloc(ctx, M * K * 4, &A);
    if (!rc) rc = gat_malloc(ctx, K * N * 4, &B);
    if (!rc) rc = gat_malloc(ctx, M * N * 4, &C);
    if (!rc) rc = gat_fill_uniform(ctx, GAT_F32, A, M * K, 0x6A7A5EED, 0, -1.0, 1.0);
    if (!rc) rc = gat_fill_uniform(ctx, GAT_F32, B, K * N, 0x6A7A5EED, 1, -1.0, 1.0);
    if (!rc) rc = gat_fill_uniform(ctx, GAT_F32, C, M * N, 0x6A7A5EED, 2, -1.0, 1.0);
    if (!rc) rc = gat_gemm(ctx, GAT_F32, 0, 0, M, N, K, 1.5, A, K, B, N, -0.5, C, N);
    std::vector<float> host(M * N);
    if (!rc) rc = gat_memcpy_d2h(ctx, host.data(), C, M * N * 4);
    if (!rc) rc = gat_sync(ctx);
    std::printf("gemm %d%s%s\n", rc, rc ? " " : "", rc ? gat_last_error() : "");
    if (rc) status = 1;
    if (!rc && argc > 1) {
        FILE* f = std::fopen(argv[1], "wb");
        if (!f || std::fwrite(host.data(), 4, host.size(), f) != host.size()) status = 1;
        if (f) std::fclose(f);
    }
    gat_free(ctx, A);
    gat_free(ctx, B);
    gat_free(ctx, C);
    gat_destroy(ctx);
    return status;
}
