#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>

#include "grouped_plan.hpp"

namespace gat {

// ---------------- arena ----------------

void DeviceArena::reset(size_t capacity) {
    free_.clear();
    live_.clear();
    used_ = 0;
    if (capacity) free_[0] = capacity;
}

size_t DeviceArena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (!bytes) bytes = 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        live_[off] = bytes;
        used_ += bytes;
        return off;
    }
    return kNone;
}

void DeviceArena::release(size_t off) {
    auto lv = live_.find(off);
    if (lv == live_.end()) return;
    size_t size = lv->second;
    live_.erase(lv);
    used_ -= size;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + size == next->first) {
        size += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += size;
            return;
        }
    }
    free_[off] = size;
}

// ---------------- trace ----------------

namespace {

struct Rng {  // SplitMix64: a fixed, portable sequence for the trace
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double log_uniform(double lo, double hi) { return std::exp(std::log(lo) + uniform() * (std::log(hi) - std::log(lo))); }
};

int64_t round_to(double v, int64_t q, int64_t lo, int64_t hi) {
    int64_t r = (int64_t)std::llround(v / (double)q) * q;
    return std::min(hi, std::max(lo, r));
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t i = (size_t)std::min<double>((double)v.size() - 1, std::ceil(p * (double)v.size()) - 1);
    return v[i];
}

}  // namespace

void Server::draw_trace() { serve_draw_trace(cfg_, reqs_, pool_elems_); }

void serve_draw_trace(const gat_serve_cfg& cfg_, std::vector<ServeReq>& reqs_, size_t& pool_elems_) {
    Rng rng{cfg_.seed};
    reqs_.clear();
    double t = 0.0;
    for (;;) {
        t += -std::log(1.0 - rng.uniform()) / cfg_.rate;
        if (t >= cfg_.duration) break;
        ServeReq r{};
        r.t_arr = t;
        r.kind = rng.uniform() < cfg_.gemm_fraction ? 0 : 1;
        if (r.kind == 0) {
            r.M = round_to(rng.log_uniform(cfg_.gemm_min, cfg_.gemm_max), 16, 16, cfg_.gemm_max);
            r.N = round_to(rng.log_uniform(cfg_.gemm_min, cfg_.gemm_max), 16, 16, cfg_.gemm_max);
            r.K = round_to(rng.log_uniform(cfg_.gemm_min, cfg_.gemm_max), 16, 16, cfg_.gemm_max);
            r.ta = (int)(rng.next() & 1);
            r.tb = (int)(rng.next() & 1);
            r.lda = r.ta ? r.M : r.K;
            r.ldb = r.tb ? r.K : r.N;
            r.out_bytes = 4 * (size_t)r.M * (size_t)r.N;
            r.foot = 4 * ((size_t)r.M * r.K + (size_t)r.K * r.N) + r.out_bytes;
        } else {
            r.M = round_to(rng.log_uniform(cfg_.gemv_min, cfg_.gemv_max), 16, 16, cfg_.gemv_max);
            r.N = round_to(rng.log_uniform(cfg_.gemv_min, cfg_.gemv_max), 16, 16, cfg_.gemv_max);
            r.K = 0;
            r.lda = r.N;
            r.ldb = 0;
            r.out_bytes = 4 * (size_t)r.M;
            r.foot = 4 * ((size_t)r.M * r.N + r.N) + r.out_bytes;
        }
        reqs_.push_back(r);
    }
    // operand placement in the resident pool: 256-byte aligned offsets
    size_t need = 1;
    for (const ServeReq& r : reqs_) {
        const size_t a = r.kind == 0 ? (size_t)(r.ta ? r.K : r.M) * r.lda : (size_t)r.M * r.lda;
        const size_t b = r.kind == 0 ? (size_t)(r.tb ? r.N : r.K) * r.ldb : (size_t)r.N;
        need = std::max(need, std::max(a, b));
    }
    pool_elems_ = need + (1u << 22);
    for (size_t i = 0; i < reqs_.size(); ++i) {
        ServeReq& r = reqs_[i];
        const size_t a = r.kind == 0 ? (size_t)(r.ta ? r.K : r.M) * r.lda : (size_t)r.M * r.lda;
        const size_t b = r.kind == 0 ? (size_t)(r.tb ? r.N : r.K) * r.ldb : (size_t)r.N;
        r.aoff = (rng.next() % ((pool_elems_ - a) / 64 + 1)) * 64;
        r.boff = (rng.next() % ((pool_elems_ - b) / 64 + 1)) * 64;
        r.keep = cfg_.keep > 0 && (i % (size_t)cfg_.keep) == 0 && (int)(i / (size_t)cfg_.keep) < cfg_.keep_max;
    }
}

// ---------------- server ----------------

Server::~Server() {
    if (!ctx_) return;
    ctx_->make_current();
    for (hipStream_t s : streams_) {
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
    }
    for (hipEvent_t e : events_all_) hipEventDestroy(e);
    for (Slot& b : slots_) {
        if (b.h_stage) hipHostFree(b.h_stage);
        if (b.d_stage) hipFree(b.d_stage);
        if (b.d_counter) hipFree(b.d_counter);
    }
    if (pool_) hipFree(pool_);
    if (arena_base_) hipFree(arena_base_);
}

MfmaCfg Server::gemm_cfg(const ServeReq& r) const { return GroupedPlan::shape_cfg(cfg_.gemm_dtype, r.shape); }

// One GEMM request alone: the single-problem kernel of the request's
// precision form with the tile configuration of its grouped tile shape, so
// the bits equal those of the grouped launch.
int Server::launch_gemm_one(const ServeReq& r, float* C, hipStream_t st) const {
    GemmArgs a;
    a.transA = r.ta;
    a.transB = r.tb;
    a.M = r.M;
    a.N = r.N;
    a.K = r.K;
    a.A = pool_ptr(r.aoff);
    a.lda = r.lda;
    a.B = pool_ptr(r.boff);
    a.ldb = r.ldb;
    a.C = C;
    a.ldc = r.N;
    if (cfg_.gemm_dtype == DT_BF16X3) return launch_gemm_mfma_bf16(DT_BF16X3, gemm_cfg(r), a, st);
    return launch_gemm_mfma_f32(gemm_cfg(r), a, st);
}

int serve_check_cfg(const gat_serve_cfg& cfg) {
    if (!(cfg.rate > 0) || !(cfg.duration > 0) || cfg.gemm_fraction < 0 || cfg.gemm_fraction > 1 ||
        cfg.streams < 1 || cfg.streams > 16 || cfg.max_batch < 1 || cfg.max_batch > 4096 || cfg.max_inflight < 1 ||
        cfg.gemm_min < 16 || cfg.gemm_max < cfg.gemm_min || cfg.gemm_max > 16384 || cfg.gemv_min < 16 ||
        cfg.gemv_max < cfg.gemv_min || cfg.gemv_max > 65536 || cfg.hbm_budget < (1u << 20) ||
        (cfg.mode != GAT_SERVE_BATCHED && cfg.mode != GAT_SERVE_LOOP) ||
        (cfg.gemm_dtype != GAT_F32 && cfg.gemm_dtype != GAT_BF16X3)) {
        set_error("serve: invalid configuration");
        return ST_INVALID;
    }
    if (cfg.rate * cfg.duration > 5e6) {
        set_error("serve: trace longer than 5e6 requests");
        return ST_INVALID;
    }
    return ST_OK;
}

int Server::create(Context& ctx, const gat_serve_cfg& cfg) {
    const int vrc = serve_check_cfg(cfg);
    if (vrc) return vrc;
    ctx_ = &ctx;
    cfg_ = cfg;
    ctx.make_current();
    draw_trace();
    for (ServeReq& r : reqs_) {
        if (r.foot > cfg_.hbm_budget) {
            set_error("serve: a request's footprint exceeds the HBM budget");
            return ST_INVALID;
        }
        if (r.kind == 0) {
            GroupedDesc d{r.ta, r.tb, r.M, r.N, r.K, 1.0, nullptr, r.lda, nullptr, r.ldb, 0.0, nullptr, r.N};
            r.shape = GroupedPlan::choose_shape(&ctx, cfg_.gemm_dtype, d);
        }
    }
    GAT_HIP_TRY(hipMalloc((void**)&pool_, pool_elems_ * sizeof(float)));
    int rc = launch_fill_uniform(DT_F32, pool_, pool_elems_, cfg_.seed, 7, -1.0, 1.0, ctx.stream);
    if (rc) return rc;
    GAT_HIP_TRY(hipMalloc(&arena_base_, cfg_.hbm_budget));
    // tile-list capacity per batched launch: the largest problem, at least 2^18
    const int64_t t = (cfg_.gemm_max + 63) / 64;
    work_cap_ = std::max<size_t>(1u << 18, (size_t)(t * t));
    streams_.resize(cfg_.streams);
    for (int i = 0; i < cfg_.streams; ++i)
        GAT_HIP_TRY(hipStreamCreateWithFlags(&streams_[i], hipStreamNonBlocking));
    if (cfg_.mode == GAT_SERVE_BATCHED) {
        // staging per batched launch in flight: GEMM descriptors followed by
        // the tile list, or GEMV descriptors; one copy moves a batch's tables
        stage_bytes_ = std::max(probs_bytes() + work_cap_ * sizeof(int4), cfg_.max_batch * sizeof(GemvProblemDev));
        slots_.resize((size_t)cfg_.streams * kSlotsPerStream);
        for (size_t i = 0; i < slots_.size(); ++i) {
            Slot& b = slots_[i];
            b.stream = (int)(i % (size_t)cfg_.streams);
            GAT_HIP_TRY(hipHostMalloc((void**)&b.h_stage, stage_bytes_));
            GAT_HIP_TRY(hipMalloc((void**)&b.d_stage, stage_bytes_));
            GAT_HIP_TRY(hipMalloc((void**)&b.d_counter, (GP_XCDS + 1) * sizeof(int)));
            GAT_HIP_TRY(hipMemsetAsync(b.d_counter, 0, (GP_XCDS + 1) * sizeof(int), ctx.stream));
        }
    }
    GAT_HIP_TRY(hipStreamSynchronize(ctx.stream));
    return ST_OK;
}

hipEvent_t Server::get_event() {
    if (!events_free_.empty()) {
        hipEvent_t e = events_free_.back();
        events_free_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    events_all_.push_back(e);
    return e;
}

int Server::launch_one(int id, int stream, Launch& L) {
    const ServeReq& r = reqs_[id];
    hipStream_t st = streams_[stream];
    int rc;
    if (r.kind == 0) {
        rc = launch_gemm_one(r, (float*)out_ptr(r), st);
    } else {
        const GemvCfg gc{256, GEMV_GROUPED_ROWS, GEMV_GROUPED_UNROLL};
        rc = launch_gemv_native(DT_F32, 0, (size_t)r.M, (size_t)r.N, 1.0, pool_ptr(r.aoff), (size_t)r.lda,
                                pool_ptr(r.boff), 0.0, out_ptr(r), nullptr, st, &gc);
    }
    if (rc) return rc;
    L.stream = stream;
    L.members.assign(1, id);
    L.ev = get_event();
    if (!L.ev) {
        set_error("serve: event creation failed");
        return ST_HIP;
    }
    GAT_HIP_TRY(hipEventRecord(L.ev, st));
    return ST_OK;
}

int Server::launch_batch(int kind, std::vector<int>& pend, int slot, Launch& L) {
    Slot& b = slots_[slot];
    hipStream_t st = streams_[b.stream];
    L.stream = b.stream;
    L.slot = slot;
    L.members.clear();
    if (kind == 0) {
        std::vector<GroupedDesc> descs;
        std::vector<int> shapes;
        size_t tiles = 0, take = 0;
        for (; take < pend.size() && (int)take < cfg_.max_batch; ++take) {
            const ServeReq& r = reqs_[pend[take]];
            const int64_t TM = gp_shape_bm(r.shape), TN = gp_shape_bn(r.shape);
            const size_t nt = (size_t)((r.M + TM - 1) / TM) * (size_t)((r.N + TN - 1) / TN);
            if (take && tiles + nt > work_cap_) break;
            tiles += nt;
            descs.push_back(GroupedDesc{r.ta, r.tb, r.M, r.N, r.K, 1.0, pool_ptr(r.aoff), r.lda, pool_ptr(r.boff),
                                        r.ldb, 0.0, out_ptr(r), r.N});
            shapes.push_back(r.shape);
            L.members.push_back(pend[take]);
        }
        std::vector<GroupedProblemDev> probs;
        std::vector<int4> work;
        std::vector<int> used;
        double flops = 0;
        GroupedQueues q{};
        int rc = GroupedPlan::build_tables(ctx_, cfg_.gemm_dtype, descs.data(), descs.size(), probs, work, q, used,
                                           flops, shapes.data());
        if (rc) return rc;
        // [descriptors | tile list] in one staging buffer, one copy
        const size_t pbytes = probs.size() * sizeof(GroupedProblemDev);
        const size_t woff = (pbytes + 15) & ~(size_t)15;
        std::memcpy(b.h_stage, probs.data(), pbytes);
        std::memcpy(b.h_stage + woff, work.data(), work.size() * sizeof(int4));
        GAT_HIP_TRY(hipMemcpyAsync(b.d_stage, b.h_stage, woff + work.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   st));
        const int grid = (int)std::max<size_t>(
            1, std::min<size_t>((size_t)grouped_wgs_per_cu(GroupedPlan::form_of(cfg_.gemm_dtype)) *
                                    (size_t)std::max(ctx_->num_cu, 1),
                                work.size()));
        rc = launch_grouped_gemm(GroupedPlan::form_of(cfg_.gemm_dtype), (const GroupedProblemDev*)b.d_stage, (const int4*)(b.d_stage + woff), q,
                                 b.d_counter, grid, st);
        if (rc) return rc;
        pend.erase(pend.begin(), pend.begin() + take);
    } else {
        const size_t take = std::min<size_t>(pend.size(), (size_t)cfg_.max_batch);
        GemvProblemDev* gv = (GemvProblemDev*)b.h_stage;
        int64_t blocks = 0;
        for (size_t i = 0; i < take; ++i) {
            const ServeReq& r = reqs_[pend[i]];
            GemvProblemDev& g = gv[i];
            g.A = pool_ptr(r.aoff);
            g.x = pool_ptr(r.boff);
            g.y = out_ptr(r);
            g.M = r.M;
            g.N = r.N;
            g.lda = r.lda;
            g.alpha = 1.f;
            g.beta = 0.f;
            g.vec = ((uintptr_t)g.A % 16 == 0) && ((uintptr_t)g.x % 16 == 0) && (g.lda % 4 == 0);
            g.blk0 = (int)blocks;
            blocks += gemv_grouped_blocks(r.M);
            L.members.push_back(pend[i]);
        }
        GAT_HIP_TRY(hipMemcpyAsync(b.d_stage, b.h_stage, take * sizeof(GemvProblemDev), hipMemcpyHostToDevice, st));
        int rc = launch_grouped_gemv((const GemvProblemDev*)b.d_stage, (int)take, blocks, st);
        if (rc) return rc;
        pend.erase(pend.begin(), pend.begin() + take);
    }
    L.ev = get_event();
    if (!L.ev) {
        set_error("serve: event creation failed");
        return ST_HIP;
    }
    GAT_HIP_TRY(hipEventRecord(L.ev, st));
    return ST_OK;
}

int Server::run(gat_serve_stats* out) {
    Context& ctx = *ctx_;
    ctx.make_current();
    arena_.reset(cfg_.hbm_budget);
    kept_.clear();
    const int n = (int)reqs_.size();
    std::deque<int> queue;
    std::vector<int> pend[2];
    std::vector<std::deque<Launch>> inflight(streams_.size());
    size_t admitted_bytes = 0, peak = 0, launches = 0, launched_reqs = 0, peak_out = 0;
    int admitted = 0, done = 0, next = 0, rr = 0, running = 0;
    // attribution: time with a launch in flight (host-observed, from issue
    // to observed completion) and time requests waited at a stopped admission
    double busy = 0.0, blocked = 0.0, tlast = 0.0;
    bool was_running = false, was_blocked = false;
    const bool batched = cfg_.mode == GAT_SERVE_BATCHED;
    const double t0 = now_s();
    while (done < n) {
        const double t = now_s() - t0;
        if (was_running) busy += t - tlast;
        if (was_blocked) blocked += t - tlast;
        tlast = t;
        bool progress = false;
        while (next < n && reqs_[next].t_arr <= t) queue.push_back(next++);
        // completions: launches on one stream finish in order
        for (size_t s = 0; s < inflight.size(); ++s) {
            while (!inflight[s].empty()) {
                Launch& L = inflight[s].front();
                const hipError_t q = hipEventQuery(L.ev);
                if (q == hipErrorNotReady) break;
                GAT_HIP_TRY(q);
                const double td = now_s() - t0;
                for (int id : L.members) {
                    ServeReq& r = reqs_[id];
                    r.t_done = td;
                    admitted_bytes -= r.foot;
                    if (r.keep) kept_.push_back(id);
                    else arena_.release(r.out_off);
                    --admitted;
                    ++done;
                }
                events_free_.push_back(L.ev);
                if (L.slot >= 0) slots_[L.slot].busy = false;
                inflight[s].pop_front();
                --running;
                progress = true;
            }
        }
        // admission: FIFO while slots and HBM budget allow
        while (!queue.empty() && admitted < cfg_.max_inflight) {
            ServeReq& r = reqs_[queue.front()];
            if (admitted_bytes + r.foot > cfg_.hbm_budget) break;
            const size_t off = arena_.alloc(r.out_bytes);
            if (off == DeviceArena::kNone) break;
            r.out_off = off;
            r.t_adm = now_s() - t0;
            admitted_bytes += r.foot;
            peak = std::max(peak, admitted_bytes);
            peak_out = std::max(peak_out, arena_.in_use());
            ++admitted;
            pend[r.kind].push_back(queue.front());
            queue.pop_front();
            progress = true;
        }
        // arrived requests left in the queue wait on the budget, max_inflight
        // or the arena
        was_blocked = !queue.empty();
        if (!queue.empty() && admitted == 0) {
            // nothing in flight will free memory: kept results fill the arena
            set_error("serve: a request cannot be admitted (kept results hold the output arena)");
            return ST_NOMEM;
        }
        // launches
        if (batched) {
            for (int kind = 1; kind >= 0; --kind) {  // GEMV first: short, latency-bound
                if (pend[kind].empty()) continue;
                if (pend[kind].size() == 1) {
                    // a batch of one is the single-problem kernel itself (same
                    // tile routine, same bits), on the least-loaded stream
                    size_t sb = 0;
                    for (size_t i = 1; i < streams_.size(); ++i)
                        if (inflight[i].size() < inflight[sb].size()) sb = i;
                    Launch L;
                    int rc = launch_one(pend[kind][0], (int)sb, L);
                    if (rc) return rc;
                    pend[kind].clear();
                    inflight[sb].push_back(std::move(L));
                    ++running;
                    ++launches;
                    ++launched_reqs;
                    progress = true;
                    continue;
                }
                // a free slot on an idle stream first (the batch overlaps the
                // running ones), else any free slot (it queues on its stream)
                int s = -1;
                for (int pass = 0; pass < 2 && s < 0; ++pass)
                    for (size_t i = 0; i < slots_.size(); ++i)
                        if (!slots_[i].busy && (pass == 1 || inflight[slots_[i].stream].empty())) {
                            s = (int)i;
                            break;
                        }
                if (s < 0) break;
                Launch L;
                int rc = launch_batch(kind, pend[kind], s, L);
                if (rc) return rc;
                slots_[s].busy = true;
                launched_reqs += L.members.size();
                ++running;
                ++launches;
                inflight[L.stream].push_back(std::move(L));
                progress = true;
            }
        } else {
            for (int kind = 1; kind >= 0; --kind) {
                for (int id : pend[kind]) {
                    Launch L;
                    int rc = launch_one(id, rr, L);
                    if (rc) return rc;
                    inflight[rr].push_back(std::move(L));
                    rr = (rr + 1) % (int)streams_.size();
                    ++running;
                    ++launches;
                    ++launched_reqs;
                }
                pend[kind].clear();
            }
        }
        (void)progress;
        was_running = running > 0;
    }
    for (hipStream_t s : streams_) GAT_HIP_TRY(hipStreamSynchronize(s));
    if (out) {
        std::memset(out, 0, sizeof(*out));
        std::vector<double> lat, lg, lv;
        double qsum = 0, last = 0;
        for (const ServeReq& r : reqs_) {
            const double ms = (r.t_done - r.t_arr) * 1e3;
            lat.push_back(ms);
            (r.kind == 0 ? lg : lv).push_back(ms);
            qsum += (r.t_adm - r.t_arr) * 1e3;
            last = std::max(last, r.t_done);
            out->flops += r.kind == 0 ? 2.0 * r.M * r.N * r.K : 2.0 * r.M * r.N;
            out->bytes += (double)r.foot;
        }
        out->requests = reqs_.size();
        out->gemm_requests = lg.size();
        out->gemv_requests = lv.size();
        out->launches = launches;
        out->kept = kept_.size();
        out->wall_s = n ? last - reqs_[0].t_arr : 0.0;
        out->tflops = out->wall_s > 0 ? out->flops / out->wall_s * 1e-12 : 0.0;
        out->req_per_s = out->wall_s > 0 ? n / out->wall_s : 0.0;
        out->p50_ms = pct(lat, 0.50);
        out->p99_ms = pct(lat, 0.99);
        out->max_ms = pct(lat, 1.0);
        out->p50_gemm_ms = pct(lg, 0.50);
        out->p99_gemm_ms = pct(lg, 0.99);
        out->p50_gemv_ms = pct(lv, 0.50);
        out->p99_gemv_ms = pct(lv, 0.99);
        out->mean_queue_ms = n ? qsum / n : 0.0;
        out->mean_batch = launches ? (double)launched_reqs / launches : 0.0;
        out->peak_footprint = (double)peak;
        out->device_busy = out->wall_s > 0 ? std::min(1.0, busy / out->wall_s) : 0.0;
        out->blocked_ms = blocked * 1e3;
        out->offered_tflops = out->flops / cfg_.duration * 1e-12;
        out->peak_output_bytes = (double)peak_out;
    }
    return ST_OK;
}

int Server::kept_get(size_t i, gat_serve_req* o) const {
    if (i >= kept_.size() || !o) {
        set_error("serve: kept index out of range");
        return ST_INVALID;
    }
    const ServeReq& r = reqs_[kept_[i]];
    o->kind = r.kind;
    o->transA = r.ta;
    o->transB = r.tb;
    o->shape = r.kind == 0 ? r.shape : -1;
    o->M = (size_t)r.M;
    o->N = (size_t)r.N;
    o->K = (size_t)r.K;
    o->A = pool_ptr(r.aoff);
    o->lda = (size_t)r.lda;
    o->B = pool_ptr(r.boff);
    o->ldb = (size_t)r.ldb;
    o->C = out_ptr(r);
    o->ldc = r.kind == 0 ? (size_t)r.N : 1;
    o->latency_ms = (r.t_done - r.t_arr) * 1e3;
    o->id = (size_t)kept_[i];
    return ST_OK;
}

int Server::check_alone(size_t* mismatches) {
    Context& ctx = *ctx_;
    ctx.make_current();
    size_t bad = 0;
    for (int id : kept_) {
        const ServeReq& r = reqs_[id];
        void* tmp = nullptr;
        GAT_HIP_TRY(hipMalloc(&tmp, r.out_bytes));
        int rc;
        if (r.kind == 0) {
            rc = launch_gemm_one(r, (float*)tmp, ctx.stream);
        } else {
            const GemvCfg gc{256, GEMV_GROUPED_ROWS, GEMV_GROUPED_UNROLL};
            rc = launch_gemv_native(DT_F32, 0, (size_t)r.M, (size_t)r.N, 1.0, pool_ptr(r.aoff), (size_t)r.lda,
                                    pool_ptr(r.boff), 0.0, tmp, nullptr, ctx.stream, &gc);
        }
        std::vector<unsigned char> h1(r.out_bytes), h2(r.out_bytes);
        if (!rc && hipMemcpyAsync(h1.data(), tmp, r.out_bytes, hipMemcpyDeviceToHost, ctx.stream) == hipSuccess &&
            hipMemcpyAsync(h2.data(), out_ptr(r), r.out_bytes, hipMemcpyDeviceToHost, ctx.stream) == hipSuccess &&
            hipStreamSynchronize(ctx.stream) == hipSuccess) {
            if (std::memcmp(h1.data(), h2.data(), r.out_bytes) != 0) ++bad;
        } else if (!rc) {
            rc = ST_HIP;
            set_error("serve: copy-back failed");
        }
        hipFree(tmp);
        if (rc) return rc;
    }
    if (mismatches) *mismatches = bad;
    return ST_OK;
}

}  // namespace gat
