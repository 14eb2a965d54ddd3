// Serving loop: an open-loop trace of GEMM / GEMV requests over
// device-resident operands, admitted by HBM footprint and continuously
// batched onto concurrent streams.
#pragma once

#include <map>
#include <vector>

#include "../../include/gatlas.h"
#include "context.hpp"
#include "kernels/gemv_grouped.hpp"
#include "kernels/grouped.hpp"

namespace gat {

// First-fit free-list allocator over one device allocation (offsets in
// bytes, 256-byte granules); adjacent free ranges are merged on release.
class DeviceArena {
public:
    static constexpr size_t kNone = ~(size_t)0;
    void reset(size_t capacity);
    size_t alloc(size_t bytes);  // kNone when no free range fits
    void release(size_t off);
    size_t in_use() const { return used_; }

private:
    std::map<size_t, size_t> free_;  // offset -> size
    std::map<size_t, size_t> live_;  // offset -> size
    size_t used_ = 0;
};

struct ServeReq {
    int kind;  // 0 GEMM, 1 GEMV
    int ta, tb, shape;
    int64_t M, N, K, lda, ldb;
    size_t aoff, boff;     // operand offsets into the resident pool (floats)
    size_t out_bytes, foot;
    size_t out_off;        // arena byte offset of C / y
    double t_arr, t_adm, t_done;
    bool keep;
};

// The request trace of a configuration: Poisson arrivals, request shapes,
// operand offsets into a resident pool of `pool_elems` floats, kept flags.
// Deterministic in cfg.seed; needs no device.
void serve_draw_trace(const gat_serve_cfg& cfg, std::vector<ServeReq>& reqs, size_t& pool_elems);
// ST_OK, or ST_INVALID with the error text set.
int serve_check_cfg(const gat_serve_cfg& cfg);

class Server {
public:
    ~Server();
    int create(Context& ctx, const gat_serve_cfg& cfg);
    int run(gat_serve_stats* st);
    size_t kept() const { return kept_.size(); }
    int kept_get(size_t i, gat_serve_req* out) const;
    int check_alone(size_t* mismatches);

private:
    struct Launch {
        int stream;
        int slot = -1;  // staging slot of a batched launch
        hipEvent_t ev;
        std::vector<int> members;
    };
    // Staging of one batched launch in flight: pinned host tables, their
    // device copy and the grouped kernel's counter pair.
    struct Slot {
        int stream = 0;
        char* h_stage = nullptr;
        char* d_stage = nullptr;
        int* d_counter = nullptr;
        bool busy = false;
    };
    static constexpr int kSlotsPerStream = 2;
    size_t probs_bytes() const {
        return ((size_t)cfg_.max_batch * sizeof(GroupedProblemDev) + 15) & ~(size_t)15;
    }
    void draw_trace();
    float* pool_ptr(size_t off) const { return pool_ + off; }
    float* out_ptr(const ServeReq& r) const { return (float*)((char*)arena_base_ + r.out_off); }
    MfmaCfg gemm_cfg(const ServeReq& r) const;
    int launch_gemm_one(const ServeReq& r, float* C, hipStream_t st) const;
    int launch_batch(int kind, std::vector<int>& pend, int slot, Launch& L);
    int launch_one(int id, int stream, Launch& L);
    hipEvent_t get_event();

    Context* ctx_ = nullptr;
    gat_serve_cfg cfg_{};
    std::vector<ServeReq> reqs_;
    std::vector<int> kept_;
    float* pool_ = nullptr;
    size_t pool_elems_ = 0;
    void* arena_base_ = nullptr;
    DeviceArena arena_;
    std::vector<hipStream_t> streams_;
    std::vector<Slot> slots_;
    size_t stage_bytes_ = 0;
    std::vector<hipEvent_t> events_free_, events_all_;
    size_t work_cap_ = 0;
};

}  // namespace gat
