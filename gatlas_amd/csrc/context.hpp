// Device context: the HIP replacement for the reference's OCLBase + OCLApp
// (one device, one in-order queue, owned buffers, device limits).
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"

namespace gat {

// Size bucket of a GEMM problem for tile selection: log2-rounded M, N, K
// (clamped), plus the transpose pair (any non-zero flag counts as 1) and the
// precision form.
struct BucketKey {
    int dtype, ta, tb, lm, ln, lk;
    bool operator<(const BucketKey& o) const {
        if (dtype != o.dtype) return dtype < o.dtype;
        if (ta != o.ta) return ta < o.ta;
        if (tb != o.tb) return tb < o.tb;
        if (lm != o.lm) return lm < o.lm;
        if (ln != o.ln) return ln < o.ln;
        return lk < o.lk;
    }
};

BucketKey bucket_of(int dtype, const GemmArgs& a);

class Context {
public:
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int max_workgroup = 1024;
    int num_cu = 0;
    size_t total_mem = 0;
    std::string arch;

    ~Context();
    int init(int dev);
    void make_current() const;
    std::string info() const;

    int alloc(size_t bytes, void** p);
    int release(void* p);
    // Grow-only scratch buffer owned by the context (GEMV partials, BF16
    // split operands, ...).  Valid until the next call that grows it.
    int workspace(size_t bytes, void** p);

    // Dispatch one GEMM (possibly strided-batched) to the kernel for dtype,
    // using `cfg` when given, else the tuned config of its size bucket, else
    // the default heuristic.
    int gemm(int dtype, const GemmArgs& a, const MfmaCfg* cfg);

    // Tuned tile choice per size bucket (filled by the autotuner).
    void set_tuned(const BucketKey& k, const MfmaCfg& c);
    bool get_tuned(const BucketKey& k, MfmaCfg* c) const;
    // Grouped-kernel tile shape per size bucket (filled by the grouped
    // tuner; a table of its own, so single-problem GEMMs keep theirs).
    void set_grouped_shape(const BucketKey& k, int shape);
    bool get_grouped_shape(const BucketKey& k, int* shape) const;
    // Forgets both tables.
    void clear_tuned();
    static MfmaCfg default_cfg(int dtype, const GemmArgs& a);
    static int num_cfgs(int dtype);
    static MfmaCfg cfg_at(int dtype, int i);
    static bool has_cfg(int dtype, const MfmaCfg& c);

private:
    int gemm_fp8(const GemmArgs& a, const MfmaCfg& c);
    int gemm_x3_split(const GemmArgs& a);
    std::unordered_map<void*, size_t> owned_;
    void* work_ = nullptr;
    size_t work_bytes_ = 0;
    std::map<BucketKey, MfmaCfg> tuned_;
    std::map<BucketKey, int> grouped_shape_;
    mutable std::mutex mu_;
};

}  // namespace gat
