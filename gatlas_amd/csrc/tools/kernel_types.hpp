// Kernel type names of the command-line tools (-T): the reference's
// float1 .. doubleimg (benchMatmul.cpp:141-160, benchMatvec.cpp:135-150,
// benchSaxpy.cpp:416-417), plus the matrix-core (…mfma) and bandwidth
// (…native) kernels of this library.
#pragma once

#include <string>

#include "../host/kernels_impl.hpp"

namespace gat {

struct KernelType {
    int dtype = DT_F32, vl = 1, variant = VAR_BUFFER;
};

inline bool parse_kernel_type(int family, const std::string& t, KernelType& k) {
    k = KernelType();
    if (t == "float1") k.vl = 1;
    else if (t == "float2") k.vl = 2;
    else if (t == "float4") k.vl = 4;
    else if (t == "double1") { k.dtype = DT_F64; k.vl = 1; }
    else if (t == "double2") { k.dtype = DT_F64; k.vl = 2; }
    else if (t == "double4") { k.dtype = DT_F64; k.vl = 4; }
    else if (t == "floatimg") { k.variant = VAR_IMAGE; k.vl = 4; }
    else if (t == "doubleimg") { k.dtype = DT_F64; k.variant = VAR_IMAGE; k.vl = 2; }
    else if (family == FAM_MATMUL && t == "floatmfma") k.variant = VAR_MFMA;
    else if (family == FAM_MATMUL && t == "doublemfma") { k.dtype = DT_F64; k.variant = VAR_MFMA; }
    else if (family == FAM_MATMUL && t == "bf16mfma") { k.dtype = DT_BF16; k.variant = VAR_MFMA; }
    else if (family == FAM_MATMUL && t == "bf16x3mfma") { k.dtype = DT_BF16X3; k.variant = VAR_MFMA; }
    else if (family == FAM_MATMUL && t == "fp8mfma") { k.dtype = DT_FP8; k.variant = VAR_MFMA; }
    else if (family != FAM_MATMUL && t == "floatnative") { k.variant = VAR_NATIVE; k.vl = 4; }
    else if (family != FAM_MATMUL && t == "doublenative") { k.dtype = DT_F64; k.variant = VAR_NATIVE; k.vl = 2; }
    else return false;
    return true;
}

}  // namespace gat
