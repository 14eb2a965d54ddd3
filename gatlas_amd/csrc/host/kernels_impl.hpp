// Concrete kernel families behind the KernelInterface boundary.
//
//   matmulbuffer{float,double}{1,2,4}, matmulimage{float4,double2}
//       reference geometry (compat engine), extra parameter digits
//       inlineMNK(2), loopOrder(6) [, globalID(2) for images] as in
//       KernelMatmulBuffer.hpp:42-49 / KernelMatmulImage.hpp:47-51;
//   matmulmfma{float,double,bf16,bf16x3,fp8}
//       matrix-core kernels: work group = waves_m x waves_n wavefronts,
//       inner blocking = tm x tn MFMA blocks per wave, extra parameter digits
//       LDS depth (bk in {8,16,32,64}) and block shape (32x32 or 16x16);
//   matvecbuffer{float,double}{1,2,4}, matvecimage{float4,double2}
//       reference geometry, digits inlineMN(2) [, globalID(2)];
//   matvecnative{float,double}
//       wavefront-reduction kernel: group size, rows per wave (blockHeight),
//       16-byte vector (vectorLength), digit unroll {1,2,4};
//   saxpybuffer{float,double}{1,2,4}, saxpyimage{float4,double2}
//       reference geometry, digits inlineMN(2) [, globalID(2)];
//   saxpynative{float,double}
//       streaming kernel (1-D): group size, vectors per work-item
//       (blockHeight), digit non-temporal {0,1}.
// The extra-parameter digits that steered OpenCL code generation (inlined
// dimensions, MAD loop order, global vs group-local ids) select among
// schedules of identical arithmetic; the compiled HIP kernels accept and
// journal them with no effect on results.
#pragma once

#include <memory>
#include <string>

#include "device_app.hpp"
#include "params.hpp"

namespace gat {

enum Family { FAM_MATMUL = 0, FAM_MATVEC = 1, FAM_SAXPY = 2 };
enum Variant { VAR_BUFFER = 0, VAR_IMAGE = 1, VAR_MFMA = 2, VAR_NATIVE = 3 };

// nullptr when the (family, dtype, vl, variant) combination does not exist.
KernelInterface* make_kernel(int family, int dtype, int vl, int variant);

// Typed access for the search drivers.
KernelBaseMatmul* as_matmul(KernelInterface* k);
KernelBaseMatvec* as_matvec(KernelInterface* k);
KernelBaseSaxpy* as_saxpy(KernelInterface* k);
int kernel_family(KernelInterface* k);
int kernel_variant(KernelInterface* k);
int kernel_dtype(KernelInterface* k);

// MFMA digit tables (LDS depth, block shape).
extern const int kMfmaBK[4];
// values of the MFMA kernels' depth digit for a precision form, and the
// depth of digit value i
size_t mfma_depths(int dtype);
int mfma_depth(int dtype, size_t i);
extern const int kMfmaBlk[2];
extern const int kMfmaWK[3];

}  // namespace gat
