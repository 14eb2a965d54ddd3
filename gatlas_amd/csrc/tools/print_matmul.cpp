// print_matmul: prints the GEMM kernel instantiation that a parameter set
// selects, with the command line of the reference's printMatmul.cpp:40-160
// (-C -T -m -n -k -g -y -x -G -a -b -v -h).  The reference prints generated
// OpenCL source; here the kernels are compiled HIP templates, so the tool
// prints the template instance, tile geometry and launch shape.  No GPU
// is needed.
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
    long M = -1, N = -1, K = -1, gs = -1, bh = -1, xp = -1;
    bool gemm = false, tA = false, tB = false, autovec = true;
    int opt;
    while ((opt = getopt(argc, argv, "habvGC:T:m:n:k:g:y:x:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0]
                          << " -T float1|float2|float4|double1|double2|double4|floatimg|doubleimg|floatmfma|"
                             "doublemfma|bf16mfma|bf16x3mfma|fp8mfma -n N [-m M -k K] [-C numKernels] -g groupSize"
                             " -y blockHeight -x extraParam [-G] [-a] [-b] [-v] [-h]\n";
                return 1;
            case 'C': packed = std::atoi(optarg); break;
            case 'T': type = optarg; break;
            case 'm': M = std::atol(optarg); break;
            case 'n': N = std::atol(optarg); break;
            case 'k': K = std::atol(optarg); break;
            case 'g': gs = std::atol(optarg); break;
            case 'y': bh = std::atol(optarg); break;
            case 'x': xp = std::atol(optarg); break;
            case 'G': gemm = true; break;
            case 'a': tA = true; break;
            case 'b': tB = true; break;
            case 'v': autovec = false; break;
        }
    }
    bool rc = true;
    KernelType kt;
    if (!parse_kernel_type(FAM_MATMUL, type, kt)) {
        std::cerr << "error: invalid kernel type of " << type << std::endl;
        rc = false;
    }
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
    const long VL = kt.variant == VAR_MFMA ? 1 : kt.vl;
    if (rc && (N % VL || M % VL || K % VL)) {
        std::cerr << "error: matrix dimensions must be multiples of " << VL << std::endl;
        rc = false;
    }
    if (gs == -1) { std::cerr << "error: must specify work item group size" << std::endl; rc = false; }
    if (bh == -1) { std::cerr << "error: must specify inner blocking height" << std::endl; rc = false; }
    if (xp == -1) { std::cerr << "error: must specify extra parameter" << std::endl; rc = false; }
    if (kt.variant != VAR_MFMA) {
        if (gs != -1 && (gs < 1 || gs > 16)) { std::cerr << "error: work item group size must be a number from 1 to 16 inclusive" << std::endl; rc = false; }
        if (tA && bh != -1 && bh % VL) { std::cerr << "error: inner blocking height must be multiple of " << VL << " when matrix A is transposed" << std::endl; rc = false; }
        if (bh != -1 && bh < VL) { std::cerr << "error: invalid inner blocking height" << std::endl; rc = false; }
    }
    if (!rc) return 1;

    std::unique_ptr<KernelInterface> kernel(make_kernel(FAM_MATMUL, kt.dtype, kt.vl, kt.variant));
    KernelBaseMatmul& k = *as_matmul(kernel.get());
    k.setUseAttrAutoVec(autovec);
    k.setPackedCalc(packed);
    k.setGeneralizedMatmul(gemm);
    k.setMatrixDimensions(M, N, K);
    k.setDataLayout(tA, tB);
    k.setWorkGroup(gs);
    k.setInnerBlocking(bh, kt.variant == VAR_MFMA ? bh : kt.vl);
    k.setExtraParameter(xp);
    if (!k.validParams()) {
        std::cerr << "error: invalid parameters (coalescedKernels, M, N, K, groupSize, blockHeight, extraParam): ("
                  << packed << ", " << M << ", " << N << ", " << K << ", " << gs << ", " << bh << ", " << xp << ")"
                  << std::endl;
        return 1;
    }
    std::cout << k;
    return 0;
}
