// bench_matmul: GEMM benchmark and autotuning search with the command line
// of the reference's benchMatmul.cpp:38-215 (same flags and journal), running
// the HIP kernels of libgatlas_hip.so on a gfx950 GPU.
//
//   -d gpu|gpuN  -j journal  -T type  -n N [-m M -k K]  [-C packed]
//   [-g groupSize [-y blockHeight [-x extraParam]]] [-t trials] [-w topN]
//   [-G] [-e] [-a] [-b] [-s] [-r] [-p] [-v] [-z] [-h]
// Types: float1 float2 float4 double1 double2 double4 floatimg doubleimg
//        (reference geometry) and floatmfma doublemfma bf16mfma bf16x3mfma
//        fp8mfma (matrix-core kernels; -g/-y give square wave/block layouts).
#include <unistd.h>

#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "../host/bench.hpp"
#include "../host/device_app.hpp"
#include "../host/kernels_impl.hpp"
#include "../host/search.hpp"

using namespace gat;

static int device_index(const std::string& d) {
    if (d == "gpu" || d == "gpu0") return 0;
    if (d.compare(0, 3, "gpu") == 0) return std::atoi(d.c_str() + 3);
    return -1;
}

int main(int argc, char* argv[]) {
    std::string device = "<unspecified>", journalFile, type = "<unspecified>";
    size_t packed = 1, trials = 1;
    long M = -1, N = -1, K = -1, gs = -1, bh = -1, xp = -1, topN = -1;
    bool gemm = false, em = false, tA = false, tB = false, toDev = false, fromDev = false, paranoid = false,
         autovec = true, debug = false;
    int opt;
    while ((opt = getopt(argc, argv, "heabsrpvzGd:j:C:T:m:n:k:g:y:x:t:w:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0]
                          << " -d gpu|gpuX -j journalFile -T float1|float2|float4|double1|double2|double4|floatimg|"
                             "doubleimg|floatmfma|doublemfma|bf16mfma|bf16x3mfma|fp8mfma -n N [-m M -k K] [-C "
                             "numKernels] [-g groupSize [-y blockHeight [-x extraParam]]] [-t numberTrials] [-w topN]"
                             " [-G] [-e] [-a] [-b] [-s] [-r] [-p] [-v] [-z] [-h]\n"
                          << "***DONE***" << std::endl;
                return 1;
            case 'd': device = optarg; break;
            case 'j': journalFile = optarg; break;
            case 'C': packed = std::atoi(optarg); break;
            case 'T': type = optarg; break;
            case 'm': M = std::atol(optarg); break;
            case 'n': N = std::atol(optarg); break;
            case 'k': K = std::atol(optarg); break;
            case 'g': gs = std::atol(optarg); break;
            case 'y': bh = std::atol(optarg); break;
            case 'x': xp = std::atol(optarg); break;
            case 't': trials = std::atoi(optarg); break;
            case 'w': topN = std::atol(optarg); break;
            case 'G': gemm = true; break;
            case 'e': em = true; break;
            case 'a': tA = true; break;
            case 'b': tB = true; break;
            case 's': toDev = true; break;
            case 'r': fromDev = true; break;
            case 'p': paranoid = true; break;
            case 'v': autovec = false; break;
            case 'z': debug = true; break;
        }
    }
    bool rc = true;
    int dtype = DT_F32, variant = VAR_BUFFER;
    size_t vl = 1, maxBH = 10, maxGS = 10;
    if (type == "float1") { vl = 1; }
    else if (type == "float2") { vl = 2; }
    else if (type == "float4") { vl = 4; }
    else if (type == "double1") { dtype = DT_F64; vl = 1; }
    else if (type == "double2") { dtype = DT_F64; vl = 2; }
    else if (type == "double4") { dtype = DT_F64; vl = 4; }
    else if (type == "floatimg") { variant = VAR_IMAGE; vl = 4; maxBH = 12; maxGS = 16; }
    else if (type == "doubleimg") { dtype = DT_F64; variant = VAR_IMAGE; vl = 2; maxBH = 12; maxGS = 16; }
    else if (type == "floatmfma") { variant = VAR_MFMA; }
    else if (type == "doublemfma") { dtype = DT_F64; variant = VAR_MFMA; }
    else if (type == "bf16mfma") { dtype = DT_BF16; variant = VAR_MFMA; }
    else if (type == "bf16x3mfma") { dtype = DT_BF16X3; variant = VAR_MFMA; }
    else if (type == "fp8mfma") { dtype = DT_FP8; variant = VAR_MFMA; }
    else { std::cerr << "error: invalid kernel type of " << type << std::endl; rc = false; }
    const int dev = device_index(device);
    if (dev < 0) { std::cerr << "error: invalid device " << device << " (gpu or gpuN)" << std::endl; rc = false; }
    if (journalFile.empty()) { std::cerr << "error: journal file must be specified" << std::endl; rc = false; }
    if (packed == 0) { std::cerr << "error: number of kernels to coalesce must be at least one" << std::endl; rc = false; }
    if (N == -1) {
        std::cerr << "error: matrix dimension N must be specified" << std::endl;
        rc = false;
    } else if (M == -1 && K == -1) {
        M = K = N;
    } else if ((M != -1) != (K != -1)) {
        std::cerr << "error: all matrix dimensions M, N and K must be specified" << std::endl;
        rc = false;
    }
    if (variant != VAR_MFMA && rc) {
        if (N % vl || M % vl || K % vl) { std::cerr << "error: matrix dimensions must be multiples of " << vl << std::endl; rc = false; }
        if (gs != -1 && (gs < 1 || gs > 16)) { std::cerr << "error: work item group size must be a number from 1 to 16 inclusive" << std::endl; rc = false; }
        if (tA && bh != -1 && bh % (long)vl) { std::cerr << "error: inner blocking height must be multiple of " << vl << " when matrix A is transposed" << std::endl; rc = false; }
        if (bh != -1 && bh < (long)vl) { std::cerr << "error: invalid inner blocking height" << std::endl; rc = false; }
    }
    if (gs == -1 && bh != -1) { std::cerr << "error: group size must be specified with block height" << std::endl; rc = false; }
    if (gs == -1 && bh == -1 && xp != -1) { std::cerr << "error: group size and block height must be specified with extra parameter" << std::endl; rc = false; }
    if (em && (gs != -1 || bh != -1 || xp != -1)) { std::cerr << "error: expectation maximization optimization will find optimal blocking" << std::endl; rc = false; }
    if (!rc) {
        std::cerr << "***DONE***" << std::endl;
        return 1;
    }

    DeviceApp app(dev);
    if (!app.ok()) {
        std::cerr << "error: no GPU " << dev << std::endl << "***DONE***" << std::endl;
        return 1;
    }
    std::unique_ptr<KernelInterface> kernel(make_kernel(FAM_MATMUL, dtype, (int)vl, variant));
    KernelBaseMatmul& k = *as_matmul(kernel.get());
    k.setUseAttrAutoVec(autovec);
    k.setPackedCalc(packed);
    Journal journal(journalFile);
    Bench bench(app, k, journal);
    LoopOpts o;
    o.numberTrials = trials;
    o.topN = topN;
    o.toDevice = toDev;
    o.fromDevice = fromDev;
    o.printDebug = debug;
    o.paranoid = paranoid;

    auto params = [&](long g, long b, long x) {
        if (variant == VAR_MFMA) {
            std::vector<size_t> fixed;
            if (g != -1 && b != -1) fixed = {(size_t)g, (size_t)g, (size_t)b, (size_t)b};
            return getParamsMatmulMfma(k, gemm, M, N, K, tA, tB, fixed, x);
        }
        return getParamsMatmul(k, app.maxWorkGroupSize(), vl, maxBH, maxGS, gemm, M, N, K, tA, tB, g, b, x);
    };

    if (!em) {
        std::vector<std::vector<size_t>> pargs = params(gs, bh, xp);
        std::vector<bool> ok(pargs.size(), true);
        std::vector<double> avg(pargs.size(), 0.0);
        journal.loadMemo();
        mainLoop(k, bench, journal, pargs, ok, avg, o);
        std::cout << "***DONE***" << std::endl;
        return 0;
    }

    SearchSpec spec;
    spec.maxWorkGroupSize = app.maxWorkGroupSize();
    spec.vl = vl;
    spec.maxBlockHeight = maxBH;
    spec.maxGroupSize = maxGS;
    spec.generalized = gemm;
    spec.tA = tA;
    spec.tB = tB;
    spec.M = M;
    spec.N = N;
    spec.K = K;
    const std::vector<size_t> best = emSearchKernel(k, bench, journal, o, spec);
    if (best.empty()) {
        std::cerr << "***DONE***" << std::endl;
        return 1;
    }
    std::cout << "***DONE***" << std::endl;
    return 0;
}
