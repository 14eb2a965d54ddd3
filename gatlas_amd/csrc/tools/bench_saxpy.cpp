// bench_saxpy: SAXPY benchmark and search with the command line of the
// reference's benchSaxpy.cpp (flags -d -j -C -T -m -n -t -w -e -s -r -p -v
// -z -h).  Types: float1 float2 float4 double1 double2 double4 floatimg
// doubleimg (reference geometry) and floatnative doublenative (streaming).
// The EM search alternates work-group shape and blocking as
// benchSaxpy.cpp:498-584 does.
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
    long M = -1, N = -1, topN = -1;
    bool em = false, toDev = false, fromDev = false, paranoid = false, debug = false;
    int opt;
    while ((opt = getopt(argc, argv, "hesrpvzd:j:C:T:m:n:t:w:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0]
                          << " -d gpu|gpuX -j journalFile -T float1|float2|float4|double1|double2|double4|floatimg|"
                             "doubleimg|floatnative|doublenative -m M [-n N] [-C numKernels] [-t numberTrials] [-w "
                             "topN] [-e] [-s] [-r] [-p] [-v] [-z] [-h]\n***DONE***"
                          << std::endl;
                return 1;
            case 'd': device = optarg; break;
            case 'j': journalFile = optarg; break;
            case 'C': packed = std::atoi(optarg); break;
            case 'T': type = optarg; break;
            case 'm': M = std::atol(optarg); break;
            case 'n': N = std::atol(optarg); break;
            case 't': trials = std::atoi(optarg); break;
            case 'w': topN = std::atol(optarg); break;
            case 'e': em = true; break;
            case 's': toDev = true; break;
            case 'r': fromDev = true; break;
            case 'p': paranoid = true; break;
            case 'v': break;
            case 'z': debug = true; break;
        }
    }
    bool rc = true;
    int dtype = DT_F32, variant = VAR_BUFFER;
    size_t vl = 1;
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
    if (M == -1) { std::cerr << "error: matrix dimension M must be specified" << std::endl; rc = false; }
    if (N == -1) N = (long)vl;
    if (rc && (M % (long)vl || N % (long)vl)) { std::cerr << "error: dimensions must be multiples of " << vl << std::endl; rc = false; }
    if (!rc) {
        std::cerr << "***DONE***" << std::endl;
        return 1;
    }
    DeviceApp app(dev);
    if (!app.ok()) {
        std::cerr << "error: no GPU " << dev << std::endl << "***DONE***" << std::endl;
        return 1;
    }
    const size_t maxBH = 16, maxGS = 256;
    std::unique_ptr<KernelInterface> kernel(make_kernel(FAM_SAXPY, dtype, (int)vl, variant));
    KernelBaseSaxpy& k = *as_saxpy(kernel.get());
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
    auto params = [&](long gH, long gW, long bh, long bw, long xp) {
        if (variant == VAR_NATIVE) return getParamsSaxpyNative(k, M, N, gH, bh, xp);
        return getParamsSaxpy(k, app.maxWorkGroupSize(), vl, maxBH, maxGS, M, N, gH, gW, bh, bw, xp);
    };
    if (!em) {
        std::vector<std::vector<size_t>> pargs = params(-1, -1, -1, -1, -1);
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
