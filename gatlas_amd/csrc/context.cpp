#include "context.hpp"

#include <cstdio>
#include <sstream>

namespace gat {

static int ilog2_clamp(int64_t v, int lo, int hi) {
    int l = 0;
    while ((int64_t(1) << (l + 1)) <= v) ++l;
    if (l < lo) l = lo;
    if (l > hi) l = hi;
    return l;
}

BucketKey bucket_of(int dtype, const GemmArgs& a) {
    return BucketKey{dtype, a.transA ? 1 : 0, a.transB ? 1 : 0, ilog2_clamp(a.M, 4, 15), ilog2_clamp(a.N, 4, 15),
                     ilog2_clamp(a.K, 4, 15)};
}

Context::~Context() {
    if (device < 0) return;
    hipSetDevice(device);
    if (stream) hipStreamSynchronize(stream);
    for (auto& kv : owned_) hipFree(kv.first);
    owned_.clear();
    if (work_) hipFree(work_);
    if (ev_start) hipEventDestroy(ev_start);
    if (ev_stop) hipEventDestroy(ev_stop);
    if (stream) hipStreamDestroy(stream);
}

int Context::init(int dev) {
    int n = 0;
    GAT_HIP_TRY(hipGetDeviceCount(&n));
    if (dev < 0 || dev >= n) {
        set_error("gat_create: device index out of range");
        return ST_INVALID;
    }
    device = dev;
    GAT_HIP_TRY(hipSetDevice(dev));
    hipDeviceProp_t p;
    GAT_HIP_TRY(hipGetDeviceProperties(&p, dev));
    max_workgroup = p.maxThreadsPerBlock;
    num_cu = p.multiProcessorCount;
    total_mem = p.totalGlobalMem;
    arch = p.gcnArchName;
    // Default-priority, non-blocking stream (the reference's per-device queue).
    GAT_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    GAT_HIP_TRY(hipEventCreate(&ev_start));
    GAT_HIP_TRY(hipEventCreate(&ev_stop));
    // device address of the allocation-free GEMV's static scratch (once per device)
    return gemv_scratch_init(dev);
}

void Context::make_current() const { hipSetDevice(device); }

std::string Context::info() const {
    std::ostringstream os;
    os << "device " << device << " arch " << arch << " CUs " << num_cu << " maxWorkGroup " << max_workgroup
       << " globalMem " << (total_mem >> 20) << " MiB";
    return os.str();
}

int Context::alloc(size_t bytes, void** p) {
    if (!p) return ST_INVALID;
    make_current();
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, bytes ? bytes : 1);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return ST_NOMEM;
    }
    std::lock_guard<std::mutex> g(mu_);
    owned_[d] = bytes;
    *p = d;
    return ST_OK;
}

int Context::release(void* p) {
    if (!p) return ST_OK;
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = owned_.find(p);
        if (it == owned_.end()) {
            set_error("gat_free: pointer not owned by this context");
            return ST_INVALID;
        }
        owned_.erase(it);
    }
    make_current();
    GAT_HIP_TRY(hipStreamSynchronize(stream));
    GAT_HIP_TRY(hipFree(p));
    return ST_OK;
}

int Context::workspace(size_t bytes, void** p) {
    if (bytes > work_bytes_) {
        make_current();
        if (work_) {
            GAT_HIP_TRY(hipStreamSynchronize(stream));
            GAT_HIP_TRY(hipFree(work_));
            work_ = nullptr;
            work_bytes_ = 0;
        }
        hipError_t e = hipMalloc(&work_, bytes);
        if (e != hipSuccess) {
            set_error(std::string("workspace hipMalloc: ") + hipGetErrorString(e));
            return ST_NOMEM;
        }
        work_bytes_ = bytes;
    }
    *p = work_;
    return ST_OK;
}

void Context::set_tuned(const BucketKey& k, const MfmaCfg& c) {
    std::lock_guard<std::mutex> g(mu_);
    tuned_[k] = c;
}

bool Context::get_tuned(const BucketKey& k, MfmaCfg* c) const {
    std::lock_guard<std::mutex> g(mu_);
    auto it = tuned_.find(k);
    if (it == tuned_.end()) return false;
    *c = it->second;
    return true;
}

void Context::set_grouped_shape(const BucketKey& k, int shape) {
    std::lock_guard<std::mutex> g(mu_);
    grouped_shape_[k] = shape;
}

bool Context::get_grouped_shape(const BucketKey& k, int* shape) const {
    std::lock_guard<std::mutex> g(mu_);
    auto it = grouped_shape_.find(k);
    if (it == grouped_shape_.end()) return false;
    *shape = it->second;
    return true;
}

void Context::clear_tuned() {
    std::lock_guard<std::mutex> g(mu_);
    tuned_.clear();
    grouped_shape_.clear();
}

int Context::num_cfgs(int dtype) {
    switch (dtype) {
        case DT_F32: return mfma_f32_num_cfgs();
        case DT_F64: return mfma_f64_num_cfgs();
        case DT_BF16:
        case DT_BF16X3: return mfma_bf16_num_cfgs(dtype);
        case DT_FP8: return mfma_fp8_num_cfgs();
    }
    return 0;
}

MfmaCfg Context::cfg_at(int dtype, int i) {
    switch (dtype) {
        case DT_F32: return mfma_f32_cfg(i);
        case DT_F64: return mfma_f64_cfg(i);
        case DT_BF16:
        case DT_BF16X3: return mfma_bf16_cfg(dtype, i);
        case DT_FP8: return mfma_fp8_cfg(i);
    }
    return MfmaCfg{0, 0, 0, 0, 0, 0};
}

bool Context::has_cfg(int dtype, const MfmaCfg& c) {
    if (c.wk != 1 && dtype != DT_F32 && dtype != DT_F64) return false;  // k-split: FP32 and FP64 tiles
    switch (dtype) {
        case DT_F32: return mfma_f32_has(c);
        case DT_F64: return mfma_f64_has(c);
        case DT_BF16:
        case DT_BF16X3: return mfma_bf16_has(dtype, c);
        case DT_FP8: return mfma_fp8_has(c);
    }
    return false;
}

// Untuned tile rule: the largest tile of the form that still gives at least
// one work-group per CU of an MI355X (256 CUs), counting batched problems;
// a problem too small to fill the chip with any tile gets the 64 x 64 tile
// of four waves (one wave when a side is below 256).  FP32 64 x 64 tiles
// that give fewer than two work-groups per CU take their k-split form (two
// wave groups share each K stage), which keeps more than one wave per SIMD.
// BF16 departs from the largest-tile rule below the 256 x 256 tile: it takes
// 64 x 64 tiles wherever 128 x 128 ones give fewer than two work-groups per
// CU.  FP64 and FP8 have rules of their own (in the function body).
// At 1024^3 FP32 this is the tile an exhaustive timing of the table picks;
// at 2048^3 and 4096^3 the table timing picks eight-wave tiles of the same
// size, within the spread of repeated timings (2048^3: 132 vs 125, 4096^3:
// 147.0 vs 146.6 TFLOP/s in one run; interleaved runs rank 2048^3 the
// other way round, 127.4 vs 126.1).
MfmaCfg Context::default_cfg(int dtype, const GemmArgs& a) {
    constexpr int64_t kFillTiles = 256;
    const int64_t batch = a.batch > 1 ? a.batch : 1;
    auto tiles = [&](const MfmaCfg& c) {
        const int64_t bm = (int64_t)c.waves_m * c.tm * c.blk, bn = (int64_t)c.waves_n * c.tn * c.blk;
        return ((a.M + bm - 1) / bm) * ((a.N + bn - 1) / bn) * batch;
    };
    auto fills = [&](const MfmaCfg& c) { return tiles(c) >= kFillTiles; };
    if (dtype == DT_F64) {
        // Eight waves of 2 x 2 16x16 blocks on a 64 x 128 tile (128 x 64
        // when A is transposed: its k-rows are then M-contiguous) wherever
        // that gives two work-groups per CU.  Interleaved timing in one
        // process, general form, against the two-group 64 x 64 tile:
        // 8192^3 NN 70.1 vs 62.6, TT 67.4 vs 63.6; 4096^3 NN 68.7 vs 61.8;
        // 2048^3 NN 65.3 vs 57.7 TFLOP/s.  Smaller problems take the 64 x 64
        // tile of two k-split wave groups staged 16 deep, whose interior
        // tiles stage straight into LDS: at 1024^3 the table timing puts it
        // at 48.5-53.5 TFLOP/s over the four transpose pairs against
        // 45.3-46.3 for four groups staged 32 deep through registers.
        const MfmaCfg wide = a.transA ? MfmaCfg{4, 2, 2, 2, 16, 16} : MfmaCfg{2, 4, 2, 2, 16, 16};
        if (tiles(wide) >= 2 * kFillTiles) return wide;
        return MfmaCfg{2, 2, 2, 2, 16, 16, 2};
    }
    const int64_t mn = a.M < a.N ? a.M : a.N;
    if (dtype == DT_FP8) {
        const MfmaCfg big{2, 4, 4, 2, 128, 32};  // 256 x 256, 8 waves
        const MfmaCfg tall{4, 2, 2, 2, 128, 32};  // 256 x 128, 8 waves
        return fills(big) ? big : fills(tall) ? tall : MfmaCfg{2, 2, 2, 2, 128, 32};
    }
    MfmaCfg big, mid, small, tiny;
    if (dtype == DT_BF16 || dtype == DT_BF16X3) {
        // 256 x 256, 8 waves; split-BF16x3 through pre-split planes
        big = dtype == DT_BF16 ? MfmaCfg{2, 4, 4, 2, 64, 32} : x3_split_cfg_value();
        // BF16: 128 x 128 staged 64 deep (interior tiles straight into LDS),
        // 3072^3 673 TFLOP/s against 579 staged 32 deep.  Split-BF16x3:
        // interior tiles take the direct-to-LDS x3 routine at any BK; BK =
        // 16 keeps the register-staged edge tiles' LDS at 48 KiB (2048^3
        // 246-260 TFLOP/s against 212-233 for BK = 32).
        mid = dtype == DT_BF16 ? MfmaCfg{2, 2, 2, 2, 64, 32} : MfmaCfg{2, 2, 2, 2, 16, 32};
        small = MfmaCfg{2, 2, 1, 1, 64, 32};
        tiny = MfmaCfg{1, 1, 2, 2, 64, 32};
    } else {
        // 256 x 256, 16 waves of 2 x 2 blocks, 32-deep stages: interior
        // tiles stage straight into LDS (sgemm_dma_tile.hpp).  Interleaved
        // timing in one process put it 0.6 % above the 8-wave tile at 8192^3
        // (NN 145.1 vs 144.2, TT 145.7 vs 144.8 TFLOP/s medians).
        big = MfmaCfg{4, 4, 2, 2, 32, 32};
        // 128 x 128, 16 waves of one block each: at 2048^3 one tile per CU,
        // and 16 waves hide the stage loads that 4 waves leave exposed
        // (interleaved timing, general form: 127.0-128.7 TFLOP/s over the
        // four transpose pairs against 119.0-121.0 for four waves of 2 x 2
        // blocks; 3072^3 102 vs 99)
        mid = MfmaCfg{4, 4, 1, 1, 32, 32};
        small = MfmaCfg{2, 2, 1, 1, 32, 32};
        tiny = MfmaCfg{1, 1, 2, 2, 32, 32};
    }
    const bool split = dtype == DT_F32;
    if (fills(big)) return big;
    // BF16: 64 x 64 tiles where 128 x 128 ones give fewer than two
    // work-groups per CU (2048^3: 531 against 296 TFLOP/s for the 128 x 128
    // tile staged 32 deep, the earlier default)
    if (dtype == DT_BF16 && tiles(mid) < 2 * kFillTiles && fills(small)) return small;
    if (fills(mid)) return mid;
    if (fills(small) || mn >= 256)
        return split && tiles(small) < 2 * kFillTiles ? MfmaCfg{2, 2, 1, 1, 32, 32, 2} : small;
    return tiny;
}

int Context::gemm(int dtype, const GemmArgs& a, const MfmaCfg* cfg) {
    MfmaCfg c;
    if (cfg) {
        c = *cfg;
    } else if (!get_tuned(bucket_of(dtype, a), &c)) {
        c = default_cfg(dtype, a);
    }
    switch (dtype) {
        case DT_F32: return launch_gemm_mfma_f32(c, a, stream);
        case DT_F64: return launch_gemm_mfma_f64(c, a, stream);
        case DT_BF16: return launch_gemm_mfma_bf16(dtype, c, a, stream);
        case DT_BF16X3:
            if (x3_split_cfg(c)) return gemm_x3_split(a);
            return launch_gemm_mfma_bf16(dtype, c, a, stream);
        case DT_FP8: return gemm_fp8(a, c);
    }
    set_error("gemm: precision form not available");
    return ST_UNSUPPORTED;
}

int Context::gemm_x3_split(const GemmArgs& a) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return ST_OK;
    if (a.K <= 0 || a.M > INT32_MAX || a.N > INT32_MAX || a.K > INT32_MAX) {
        // nothing to split: the register-staged form handles the epilogue
        return launch_gemm_mfma_bf16(DT_BF16X3, MfmaCfg{2, 4, 4, 2, 16, 32}, a, stream);
    }
    const X3SplitDims d = x3_split_dims(a.M, a.N, a.K);
    void* w = nullptr;
    int rc = workspace(d.bytes_a + d.bytes_b, &w);
    if (rc) return rc;
    unsigned char* base = (unsigned char*)w;
    // op(A)[i][k] = A[i*lda + k] or A[k*lda + i]; op(B)^T[j][k] = B[k*ldb + j] or B[j*ldb + k]
    const int64_t ars = a.transA ? 1 : a.lda, acs = a.transA ? a.lda : 1;
    const int64_t brs = a.transB ? a.ldb : 1, bcs = a.transB ? 1 : a.ldb;
    for (int b = 0; b < a.batch; ++b) {
        const float* A = (const float*)a.A + (int64_t)b * a.strideA;
        const float* B = (const float*)a.B + (int64_t)b * a.strideB;
        float* C = (float*)a.C + (int64_t)b * a.strideC;
        rc = launch_x3_split(A, ars, acs, (int)a.M, (int)a.K, d.mpad, d.kpad, base, stream);
        if (!rc) rc = launch_x3_split(B, brs, bcs, (int)a.N, (int)a.K, d.npad, d.kpad, base + d.bytes_a, stream);
        if (!rc)
            rc = launch_x3_split_gemm(d, base, base + d.bytes_a, C, (int)a.M, (int)a.N, a.ldc, (float)a.alpha,
                                      (float)a.beta, stream);
        if (rc) return rc;
    }
    return ST_OK;
}

int Context::gemm_fp8(const GemmArgs& a, const MfmaCfg& c) {
    if (a.scaleA && a.scaleB) return launch_gemm_fp8(c, a, stream);
    // FP32 operands: quantise op(A) and op(B)^T into the workspace first.
    if (a.batch != 1) {
        set_error("fp8 gemm from FP32 operands: batch must be 1");
        return ST_INVALID;
    }
    const size_t M = a.M, N = a.N, K = a.K;
    const size_t kb = (K + 127) / 128, mb = (M + 127) / 128, nb = (N + 127) / 128;
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t offB = up(M * K), offSA = offB + up(N * K), offSB = offSA + up(mb * kb * 4);
    void* w = nullptr;
    int rc = workspace(offSB + up(nb * kb * 4), &w);
    if (rc) return rc;
    unsigned char* base = (unsigned char*)w;
    float* sA = (float*)(base + offSA);
    float* sB = (float*)(base + offSB);
    // op(A)[i][k]: A[i*lda + k] or A[k*lda + i]; op(B)^T[j][k] = op(B)[k][j]
    rc = launch_quantize_fp8((const float*)a.A, a.transA ? 1 : a.lda, a.transA ? a.lda : 1, (int)M, (int)K, base,
                             sA, stream);
    if (rc) return rc;
    rc = launch_quantize_fp8((const float*)a.B, a.transB ? a.ldb : 1, a.transB ? 1 : a.ldb, (int)N, (int)K,
                             base + offB, sB, stream);
    if (rc) return rc;
    GemmArgs q = a;
    q.A = base;
    q.B = base + offB;
    q.transA = 0;
    q.transB = 1;
    q.lda = a.K;
    q.ldb = a.K;
    q.scaleA = sA;
    q.scaleB = sB;
    return launch_gemm_fp8(c, q, stream);
}

}  // namespace gat
