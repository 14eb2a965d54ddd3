// print_saxpy: prints the SAXPY kernel instantiation that a parameter set
// selects, with the command line of the reference's printSaxpy.cpp:40-140
// (-C -T -m -n -G -Y -g -y -x -v -h: -Y/-G work-group height/width, -y/-g
// blocking height/width).  No GPU is needed.
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
    long M = -1, N = -1, gH = -1, gW = -1, bh = -1, bw = -1, xp = -1;
    bool autovec = true;
    int opt;
    while ((opt = getopt(argc, argv, "hvC:T:m:n:G:Y:g:y:x:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0]
                          << " -T float1|float2|float4|double1|double2|double4|floatimg|doubleimg|floatnative|"
                             "doublenative -m M [-n N] [-C numKernels] -Y groupHeight -G groupWidth -y blockHeight"
                             " -g blockWidth -x extraParam [-v] [-h]\n";
                return 1;
            case 'C': packed = std::atoi(optarg); break;
            case 'T': type = optarg; break;
            case 'm': M = std::atol(optarg); break;
            case 'n': N = std::atol(optarg); break;
            case 'Y': gH = std::atol(optarg); break;
            case 'G': gW = std::atol(optarg); break;
            case 'y': bh = std::atol(optarg); break;
            case 'g': bw = std::atol(optarg); break;
            case 'x': xp = std::atol(optarg); break;
            case 'v': autovec = false; break;
        }
    }
    bool rc = true;
    KernelType kt;
    if (!parse_kernel_type(FAM_SAXPY, type, kt)) {
        std::cerr << "error: invalid kernel type of " << type << std::endl;
        rc = false;
    }
    if (packed == 0) { std::cerr << "error: number of kernels to coalesce must be at least one" << std::endl; rc = false; }
    const long VL = kt.vl;
    if (M == -1) {
        std::cerr << "error: matrix dimension M must be specified" << std::endl;
        rc = false;
    } else if (N == -1) {
        N = VL;
    }
    if (rc && M % VL) { std::cerr << "error: matrix dimension M must be multiple of " << VL << std::endl; rc = false; }
    if (rc && N % VL) { std::cerr << "error: matrix dimension N must be multiple of " << VL << std::endl; rc = false; }
    if (gH == -1) { std::cerr << "error: must specify work item group height" << std::endl; rc = false; }
    if (gW == -1) { std::cerr << "error: must specify work item group width" << std::endl; rc = false; }
    if (bh == -1) { std::cerr << "error: must specify inner blocking height" << std::endl; rc = false; }
    if (bw == -1) { std::cerr << "error: must specify inner blocking width" << std::endl; rc = false; }
    if (xp == -1) { std::cerr << "error: must specify extra parameter" << std::endl; rc = false; }
    if (!rc) return 1;

    std::unique_ptr<KernelInterface> kernel(make_kernel(FAM_SAXPY, kt.dtype, kt.vl, kt.variant));
    KernelBaseSaxpy& k = *as_saxpy(kernel.get());
    k.setUseAttrAutoVec(autovec);
    k.setPackedCalc(packed);
    k.setSaxpyDimensions(M, N);
    k.setVectorLength(kt.vl);
    k.setWorkGroup(gH, gW);
    k.setInnerBlocking(bh, bw);
    k.setExtraParameter(xp);
    if (!k.validParams()) {
        std::cerr << "error: invalid parameters (coalescedKernels, M, N, groupHeight, groupWidth, blockHeight, "
                     "blockWidth, extraParam): ("
                  << packed << ", " << M << ", " << N << ", " << gH << ", " << gW << ", " << bh << ", " << bw << ", "
                  << xp << ")" << std::endl;
        return 1;
    }
    std::cout << k;
    return 0;
}
