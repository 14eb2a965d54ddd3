// Shared definitions for the MI355X (gfx950) GEMM / GEMV / SAXPY library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

namespace gat {

// Scalar / precision forms of the hot path.  F32 and F64 are the
// reference's two scalar types (benchMatmul.cpp:141-160 "float*", "double*");
// the rest are MFMA-only input forms that all accumulate in FP32.
enum DType : int {
    DT_F32 = 0,      // FP32 in, FP32 out
    DT_F64 = 1,      // FP64 in, FP64 out
    DT_BF16 = 2,     // BF16 in (device arrays of bf16), FP32 accumulate/out
    DT_BF16X3 = 3,   // FP32 in, split into BF16 hi+lo, 3 MFMA products, FP32 out
    DT_FP8 = 4,      // FP8 e4m3 in with per-128x128-block FP32 scales, FP32 out
};

inline size_t dtype_in_bytes(int dt) {
    switch (dt) {
        case DT_F32: return 4;
        case DT_F64: return 8;
        case DT_BF16: return 2;
        case DT_BF16X3: return 4;
        case DT_FP8: return 1;
    }
    return 0;
}

inline size_t dtype_out_bytes(int dt) { return dt == DT_F64 ? 8 : 4; }

// Status codes of the C ABI (0 ok, negative error).
enum Status : int {
    ST_OK = 0,
    ST_INVALID = -1,     // bad argument or unsupported shape/config
    ST_HIP = -2,         // HIP runtime error
    ST_NOMEM = -3,       // allocation failed / admission refused
    ST_UNSUPPORTED = -4, // config not instantiated
};

}  // namespace gat

#define GAT_HIP_TRY(expr)                                                   \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) {                                             \
            ::gat::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return ::gat::ST_HIP;                                           \
        }                                                                   \
    } while (0)

namespace gat {
// Thread-local last error text (also mirrored per context in the C ABI).
void set_error(const std::string& msg);
const char* last_error();
// Raises `kernel`'s dynamic LDS limit to `bytes` on the current device; the
// attribute call is made once per (kernel, device), later calls are free.
hipError_t allow_dynamic_lds(const void* kernel, size_t bytes);
}  // namespace gat
