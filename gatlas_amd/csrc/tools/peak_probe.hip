// peak_probe: sustained matrix-core rates (FP32 32x32x2 and BF16 32x32x16
// MFMA loops with no memory traffic) and sustained HBM read / copy bandwidth
// of plain streaming kernels -- the practical ceilings kernels are judged by.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

__global__ __launch_bounds__(256) void mfma_f32_loop(float* out, int iters) {
    f16v acc[8];
    for (int i = 0; i < 8; ++i) for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;
    float a = (float)(threadIdx.x & 7) * 0.001f, b = 0.5f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[i], 0, 0, 0);
    }
    float s = 0.f;
    for (int i = 0; i < 8; ++i) for (int r = 0; r < 16; ++r) s += acc[i][r];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void mfma_bf16_loop(float* out, int iters) {
    f16v acc[8];
    for (int i = 0; i < 8; ++i) for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) { a[e] = (__bf16)((threadIdx.x & 7) * 0.001f); b[e] = (__bf16)0.5f; }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[i], 0, 0, 0);
    }
    float s = 0.f;
    for (int i = 0; i < 8; ++i) for (int r = 0; r < 16; ++r) s += acc[i][r];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// read-only reduction: each thread sums float4s grid-stride
__global__ __launch_bounds__(256) void read_kernel(const f4v* __restrict__ x, size_t n4, float* out) {
    f4v s = {0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
        s += __builtin_nontemporal_load(x + i);
    out[blockIdx.x * 256 + threadIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(256) void copy_kernel(const f4v* __restrict__ x, f4v* __restrict__ y, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
        __builtin_nontemporal_store(__builtin_nontemporal_load(x + i), y + i);
}

int main() {
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    printf("device %s CUs %d clock %d MHz\n", p.gcnArchName, cus, p.clockRate / 1000);
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int grid = cus * 8;  // 8 work-groups of 4 waves per CU = 8 waves per SIMD requested
    float* out;
    CHECK(hipMalloc(&out, (size_t)grid * 256 * sizeof(float) * 16));
    for (int kind = 0; kind < 2; ++kind) {
        for (int g : {cus * 2, cus * 4, cus * 8}) {
            const int iters = 2000;
            if (kind == 0) hipLaunchKernelGGL(mfma_f32_loop, dim3(g), dim3(256), 0, 0, out, 10);
            else hipLaunchKernelGGL(mfma_bf16_loop, dim3(g), dim3(256), 0, 0, out, 10);
            CHECK(hipDeviceSynchronize());
            CHECK(hipEventRecord(e0, 0));
            if (kind == 0) hipLaunchKernelGGL(mfma_f32_loop, dim3(g), dim3(256), 0, 0, out, iters);
            else hipLaunchKernelGGL(mfma_bf16_loop, dim3(g), dim3(256), 0, 0, out, iters);
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            const double flop_per = kind == 0 ? 2.0 * 32 * 32 * 2 : 2.0 * 32 * 32 * 16;
            const double flops = flop_per * 8.0 * iters * (double)g * 4;  // per wave: 8 mfma per iter; 4 waves per WG
            printf("%s grid %d: %.3f ms, %.1f TFLOP/s\n", kind == 0 ? "mfma_f32_32x32x2" : "mfma_bf16_32x32x16", g, ms,
                   flops / (ms * 1e-3) / 1e12);
        }
    }
    const size_t n = (size_t)1 << 28;  // floats
    float *x, *y;
    CHECK(hipMalloc(&x, n * 4));
    CHECK(hipMalloc(&y, n * 4));
    CHECK(hipMemset(x, 0, n * 4));
    for (int g : {cus * 4, cus * 8, cus * 16, cus * 32}) {
        for (int rep = 0; rep < 2; ++rep) {
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(read_kernel, dim3(g), dim3(256), 0, 0, (const f4v*)x, n / 4, out);
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (rep) printf("read  grid %d: %.3f ms %.1f GB/s\n", g, ms, n * 4.0 / (ms * 1e-3) / 1e9);
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, 0, (const f4v*)x, (f4v*)y, n / 4);
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (rep) printf("copy  grid %d: %.3f ms %.1f GB/s\n", g, ms, n * 8.0 / (ms * 1e-3) / 1e9);
        }
    }
    return 0;
}
