// print_matvec: prints the GEMV kernel instantiation that a parameter set
// selects, with the command line of the reference's printMatvec.cpp:40-150
// (-C -T -m -n -g -y -x -G -a -v -h).  Types float1 .. doubleimg are the
// reference-geometry kernels, floatnative / doublenative the wavefront-
// reduction kernels.  No GPU is needed.
#include <unistd.h>

#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "kernel_types.hpp"

using namespace gat;

int main(int argc, char* argv[]) {
    std::string type = "<unspecified>";
    size_t packed = 1;
    long M = -1, N = -1, gs = -1, bh = -1, xp = -1;
    bool gemv = false, tA = false, autovec = true;
    int opt;
    while ((opt = getopt(argc, argv, "havGC:T:m:n:g:y:x:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0]
                          << " -T float1|float2|float4|double1|double2|double4|floatimg|doubleimg|floatnative|"
                             "doublenative -n N [-m M] [-C numKernels] -g groupSize -y blockHeight -x extraParam"
                             " [-G] [-a] [-v] [-h]\n";
                return 1;
            case 'C': packed = std::atoi(optarg); break;
            case 'T': type = optarg; break;
            case 'm': M = std::atol(optarg); break;
            case 'n': N = std::atol(optarg); break;
            case 'g': gs = std::atol(optarg); break;
            case 'y': bh = std::atol(optarg); break;
            case 'x': xp = std::atol(optarg); break;
            case 'G': gemv = true; break;
            case 'a': tA = true; break;
            case 'v': autovec = false; break;
        }
    }
    bool rc = true;
    KernelType kt;
    if (!parse_kernel_type(FAM_MATVEC, type, kt)) {
        std::cerr << "error: invalid kernel type of " << type << std::endl;
        rc = false;
    }
    if (packed == 0) { std::cerr << "error: number of kernels to coalesce must be at least one" << std::endl; rc = false; }
    if (N == -1) {
        std::cerr << "error: matrix dimension N must be specified" << std::endl;
        rc = false;
    } else if (M == -1) {
        M = N;
    }
    const long VL = kt.vl;
    if (rc && N % VL) { std::cerr << "error: matrix dimension N must be multiple of " << VL << std::endl; rc = false; }
    if (rc && M % VL) { std::cerr << "error: matrix dimension M must be multiple of " << VL << std::endl; rc = false; }
    if (gs == -1) { std::cerr << "error: must specify work item group size" << std::endl; rc = false; }
    if (bh == -1) { std::cerr << "error: must specify inner blocking height" << std::endl; rc = false; }
    if (xp == -1) { std::cerr << "error: must specify extra parameter" << std::endl; rc = false; }
    if (!rc) return 1;

    std::unique_ptr<KernelInterface> kernel(make_kernel(FAM_MATVEC, kt.dtype, kt.vl, kt.variant));
    KernelBaseMatvec& k = *as_matvec(kernel.get());
    k.setUseAttrAutoVec(autovec);
    k.setPackedCalc(packed);
    k.setGeneralizedMatvec(gemv);
    k.setMatrixDimensions(M, N);
    k.setDataLayout(tA);
    k.setWorkGroup(gs);
    k.setInnerBlocking(bh, kt.vl);
    k.setExtraParameter(xp);
    if (!k.validParams()) {
        std::cerr << "error: invalid parameters (coalescedKernels, M, N, groupSize, blockHeight, extraParam): ("
                  << packed << ", " << M << ", " << N << ", " << gs << ", " << bh << ", " << xp << ")" << std::endl;
        return 1;
    }
    std::cout << k;
    return 0;
}
