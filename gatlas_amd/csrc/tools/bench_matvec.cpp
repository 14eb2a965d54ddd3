// bench_matvec: GEMV benchmark and search with the command line of the
// reference's benchMatvec.cpp (flags -d -j -C -T -m -n -g -y -x -t -w -G -e
// -a -s -r -p -v -z -h).  Types: float1 float2 float4 double1 double2 double4
// floatimg doubleimg (reference geometry) and floatnative doublenative
// (wavefront-reduction kernels).
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

int main(int argc, char* argv[]) {
    std::string device = "<unspecified>", journalFile, type = "<unspecified>";
    size_t packed = 1, trials = 1;
    long M = -1, N = -1, gs = -1, bh = -1, xp = -1, topN = -1;
    bool gemv = false, em = false, tA = false, toDev = false, fromDev = false, paranoid = false, debug = false;
    int opt;
    while ((opt = getopt(argc, argv, "heasrpvzGd:j:C:T:m:n:g:y:x:t:w:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0]
                          << " -d gpu|gpuX -j journalFile -T float1|float2|float4|double1|double2|double4|floatimg|"
                             "doubleimg|floatnative|doublenative -n N [-m M] [-C numKernels] [-g groupSize [-y "
                             "blockHeight [-x extraParam]]] [-t numberTrials] [-w topN] [-G] [-e] [-a] [-s] [-r] [-p]"
                             " [-v] [-z] [-h]\n***DONE***"
                          << std::endl;
                return 1;
            case 'd': device = optarg; break;
            case 'j': journalFile = optarg; break;
            case 'C': packed = std::atoi(optarg); break;
            case 'T': type = optarg; break;
            case 'm': M = std::atol(optarg); break;
            case 'n': N = std::atol(optarg); break;
            case 'g': gs = std::atol(optarg); break;
            case 'y': bh = std::atol(optarg); break;
            case 'x': xp = std::atol(optarg); break;
            case 't': trials = std::atoi(optarg); break;
            case 'w': topN = std::atol(optarg); break;
            case 'G': gemv = true; break;
            case 'e': em = true; break;
            case 'a': tA = true; break;
            case 's': toDev = true; break;
            case 'r': fromDev = true; break;
            case 'p': paranoid = true; break;
            case 'v': break;
            case 'z': debug = true; break;
        }
    }
    bool rc = true;
    int dtype = DT_F32, variant = VAR_BUFFER;
    size_t vl = 1, maxBH = 64, maxGS = 256;
    if (type == "float1") vl = 1;
    else if (type == "float2") vl = 2;
    else if (type == "float4") vl = 4;
    else if (type == "double1") { dtype = DT_F64; vl = 1; }
    else if (type == "double2") { dtype = DT_F64; vl = 2; }
    else if (type == "double4") { dtype = DT_F64; vl = 4; }
    else if (type == "floatimg") { variant = VAR_IMAGE; vl = 4; }
    else if (type == "doubleimg") { dtype = DT_F64; variant = VAR_IMAGE; vl = 2; }
    else if (type == "floatnative") { variant = VAR_NATIVE; vl = 4; }
    else if (type == "doublenative") { dtype = DT_F64; variant = VAR_NATIVE; vl = 2; }
    else { std::cerr << "error: invalid kernel type of " << type << std::endl; rc = false; }
    int dev = -1;
    if (device == "gpu" || device == "gpu0") dev = 0;
    else if (device.compare(0, 3, "gpu") == 0) dev = std::atoi(device.c_str() + 3);
    if (dev < 0) { std::cerr << "error: invalid device " << device << std::endl; rc = false; }
    if (journalFile.empty()) { std::cerr << "error: journal file must be specified" << std::endl; rc = false; }
    if (N == -1) { std::cerr << "error: matrix dimension N must be specified" << std::endl; rc = false; }
    if (M == -1) M = N;
    if (rc && (M % (long)vl || N % (long)vl)) { std::cerr << "error: dimensions must be multiples of " << vl << std::endl; rc = false; }
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
    std::unique_ptr<KernelInterface> kernel(make_kernel(FAM_MATVEC, dtype, (int)vl, variant));
    KernelBaseMatvec& k = *as_matvec(kernel.get());
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
        if (variant == VAR_NATIVE) return getParamsMatvecNative(k, gemv, M, N, tA, g, b, x);
        return getParamsMatvec(k, app.maxWorkGroupSize(), vl, maxBH, maxGS, gemv, M, N, tA, g, b, x);
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
    spec.generalized = gemv;
    spec.tA = tA;
    spec.M = M;
    spec.N = N;
    const std::vector<size_t> best = emSearchKernel(k, bench, journal, o, spec);
    if (best.empty()) {
        std::cerr << "***DONE***" << std::endl;
        return 1;
    }
    std::cout << "***DONE***" << std::endl;
    return 0;
}
