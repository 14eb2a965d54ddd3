#include "bench.hpp"

#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>

#include "device_app.hpp"

namespace gat {

static thread_local bool g_quiet = false;

std::ostream& report() {
    static std::ostream null_stream(nullptr);  // no buffer: every write is dropped
    return g_quiet ? null_stream : std::cout;
}

QuietReports::QuietReports(bool on) : prev_(g_quiet) {
    if (on) g_quiet = true;
}
QuietReports::~QuietReports() { g_quiet = prev_; }

// ------------------------------------------------------------- Journal
std::string Journal::key(const KernelInterface& k, const std::vector<size_t>& params) {
    std::ostringstream s;
    s << k.kernelName() << "_";
    for (size_t p : params) s << p << "_";
    return s.str();
}

bool Journal::loadMemo() {
    std::ifstream in(file_.c_str());
    if (!in.is_open()) return false;
    state_.clear();
    times_.clear();
    std::string k;
    int v;
    while (in >> k >> v) {
        if (v < 0) state_[k] = v;
        else times_[k].push_back((size_t)v);
    }
    return true;
}

int Journal::purgeMemo(bool deleteTimes) {
    std::ofstream out(file_.c_str());
    if (!out.is_open()) return -1;
    int bad = 0;
    for (const auto& kv : state_) {
        auto it = times_.find(kv.first);
        if (it == times_.end()) {
            out << kv.first << "\t" << kv.second << "\n";
            ++bad;
        } else if (!deleteTimes) {
            out << kv.first << "\t" << kv.second << "\n";
            for (size_t t : it->second) out << kv.first << "\t" << t << "\n";
        }
    }
    return bad;
}

int Journal::memoRunState(const KernelInterface& k, const std::vector<size_t>& params) const {
    auto it = state_.find(key(k, params));
    return it == state_.end() ? MISSING : it->second;
}

int Journal::memoTime(const KernelInterface& k, const std::vector<size_t>& params, size_t trial) const {
    return memoTime(key(k, params), trial);
}

int Journal::memoTime(const std::string& key, size_t trial) const {
    auto it = times_.find(key);
    if (it == times_.end() || trial >= it->second.size()) return -1;
    return (int)it->second[trial];
}

bool Journal::takeMemo(const KernelInterface& k, const std::vector<size_t>& params, int value) const {
    return takeMemo(key(k, params), value);
}

bool Journal::takeMemo(const std::string& key, int value) const {
    std::ofstream out(file_.c_str(), std::ios::app);
    if (!out.is_open()) return false;
    out << key << "\t" << value << std::endl;  // flushed: the record must survive a crash
    return true;
}

// --------------------------------------------------------------- Bench
size_t Bench::run(size_t numTrials, const std::vector<size_t>& args, bool toDevice, bool fromDevice,
                  bool printDebug) {
    k_.setParams(args);
    if (print_ && printDebug) std::cerr << k_ << std::endl;
    if (journal_) journal_->takeMemo(k_, args, Journal::BUILD_IN_PROGRESS);
    if (stageHook) stageHook("build");
    if (print_) std::cerr << "rebuilding kernel...";
    if (!k_.build(app_)) return 0;
    if (print_) std::cerr << " done\t";
    if (journal_) journal_->takeMemo(k_, args, Journal::BUILD_OK);
    if (stageHook) stageHook("setargs");
    if (!toDevice && !k_.setArgs(app_, toDevice)) return 0;
    if (journal_) journal_->takeMemo(k_, args, Journal::RUN_IN_PROGRESS);
    if (stageHook) stageHook("run");

    const auto t0 = std::chrono::steady_clock::now();
    if (toDevice && !k_.setArgs(app_, toDevice)) return 0;
    for (size_t i = 0; i < numTrials * launches_; ++i) {
        if (!k_.enqueue(app_)) {
            if (print_) std::cerr << "error: enqueue kernel " << i << std::endl;
            app_.wait();
            return 0;
        }
    }
    if (!app_.wait()) {
        if (print_) std::cerr << "error: waiting for all kernels" << std::endl;
        return 0;
    }
    if (fromDevice && !k_.syncOutput(app_)) {
        if (print_) std::cerr << "error: read output data from device" << std::endl;
        return 0;
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (!fromDevice && !k_.syncOutput(app_)) {
        if (print_) std::cerr << "error: read output data from device" << std::endl;
        return 0;
    }
    size_t us = (size_t)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / launches_;
    if (us == 0) us = 1;  // 0 is reserved for "failed check"

    const bool isOk = k_.checkOutput(app_, printDebug);
    if (!isOk && print_) std::cerr << "fail";
    app_.wait();
    if (journal_) {
        journal_->takeMemo(k_, args, Journal::RUN_OK);
        journal_->takeMemo(k_, args, isOk ? (int)us : 0);
    }
    return isOk ? us : 0;
}

// ------------------------------------------------------------- AppUtil
namespace AppUtil {

void benchInit(const std::vector<std::vector<size_t>>& pargs, std::vector<bool>& ok, std::vector<size_t>& time,
               std::vector<size_t>& flops, std::vector<double>& avg, std::vector<double>& var,
               std::vector<std::vector<size_t>>& extra) {
    const size_t n = pargs.size();
    ok.assign(n, true);
    time.assign(n, 0);
    flops.assign(n, 0);
    avg.assign(n, 0.0);
    var.assign(n, 0.0);
    extra.assign(n, std::vector<size_t>());
}

static void print_args(const std::vector<size_t>& args, const std::vector<size_t>& extra) {
    for (size_t i = 0; i < args.size(); ++i) report() << args[i] << (i + 1 < args.size() ? " " : "");
    report() << "\t(";
    for (size_t i = 0; i < extra.size(); ++i) report() << extra[i] << (i + 1 < extra.size() ? " " : "");
    report() << ")" << std::endl;
}

size_t benchLoop(size_t trialNumber, KernelInterface& kernel, Bench& bench, Journal* journal,
                 const std::vector<std::vector<size_t>>& pargs, std::vector<bool>& ok, std::vector<size_t>& time,
                 std::vector<size_t>& flops, std::vector<double>& avg, std::vector<double>& var,
                 std::vector<std::vector<size_t>>& extra, bool toDevice, bool fromDevice, bool dummyRun,
                 bool printDebug) {
    const bool printStatus = bench.printStatus();
    size_t good = 0;
    bool needDummy = dummyRun;
    for (size_t j = 0; j < pargs.size(); ++j) {
        const std::vector<size_t>& args = pargs[j];
        if (!ok[j]) continue;
        const int memoState = journal ? journal->memoRunState(kernel, args) : Journal::MISSING;
        if (needDummy && memoState == Journal::MISSING) {
            report() << "[dummy run] ";
            bench.run(1, args, toDevice, fromDevice, printDebug);
            report() << std::endl;
            needDummy = false;
        }
        if (printStatus) report() << "[trial " << trialNumber << "] ";
        size_t us;
        if (memoState == Journal::MISSING) {
            us = bench.run(1, args, toDevice, fromDevice, printDebug);
        } else if (memoState == Journal::RUN_OK) {
            const int memo = journal->memoTime(kernel, args, trialNumber);
            us = (memo == -1) ? bench.run(1, args, toDevice, fromDevice, printDebug) : (size_t)memo;
            kernel.setParams(args);
        } else {
            // crashed while building or setting arguments, or hung while running
            if (printStatus) report() << "bad";
            us = 0;
            kernel.setParams(args);
        }
        extra[j] = kernel.extraParamDetail();
        if (us == 0) {
            if (printStatus || !journal) {
                report() << "\t";
                print_args(args, extra[j]);
            }
            ok[j] = false;
            continue;
        }
        ++good;
        const size_t nf = kernel.numberFlops();
        time[j] += us;
        flops[j] += nf;
        const double g = (double)nf / (double)us / 1000.0;  // GFLOPS
        if (trialNumber == 0) {
            avg[j] = g;
        } else {
            const double delta = g - avg[j];
            avg[j] += (1.0 / (trialNumber + 1)) * delta;
            var[j] += ((double)trialNumber / (trialNumber + 1)) * delta * delta;
        }
        if (printStatus) {
            report() << us << "\t";
            print_args(args, extra[j]);
        }
    }
    return good;
}

void markBench(size_t topN, std::vector<bool>& ok, const std::vector<size_t>& time) {
    std::set<size_t> times;
    for (size_t i = 0; i < ok.size(); ++i)
        if (ok[i]) times.insert(time[i]);
    std::set<size_t> top;
    size_t count = 0;
    for (size_t t : times) {
        if (count++ >= topN) break;
        top.insert(t);
    }
    for (size_t i = 0; i < ok.size(); ++i)
        if (ok[i] && !top.count(time[i])) ok[i] = false;
}

void markBenchAvg(size_t topN, std::vector<bool>& ok, const std::vector<double>& avg) {
    std::set<double> avgs;
    for (size_t i = 0; i < ok.size(); ++i)
        if (ok[i]) avgs.insert(avg[i]);
    std::set<double> top;
    size_t count = 0;
    for (auto it = avgs.rbegin(); it != avgs.rend(); ++it) {
        if (count++ >= topN) break;
        top.insert(*it);
    }
    for (size_t i = 0; i < ok.size(); ++i)
        if (ok[i] && !top.count(avg[i])) ok[i] = false;
}

int rankBench(size_t nthPlace, const std::vector<bool>& ok, const std::vector<double>& avg) {
    std::map<double, size_t> byAvg;  // equal averages: the later index overwrites
    for (size_t i = 0; i < ok.size(); ++i)
        if (ok[i]) byAvg[avg[i]] = i;
    size_t count = 0;
    for (auto it = byAvg.rbegin(); it != byAvg.rend(); ++it)
        if (nthPlace == count++) return (int)it->second;
    return -1;
}

void printBench(size_t numberTrials, const std::vector<std::vector<size_t>>& pargs, const std::vector<bool>& ok,
                const std::vector<size_t>& time, const std::vector<double>& avg, const std::vector<double>& var,
                const std::vector<std::vector<size_t>>& extra) {
    std::map<size_t, size_t> byTime;  // equal times: the later index overwrites
    for (size_t i = 0; i < pargs.size(); ++i)
        if (ok[i]) byTime[time[i]] = i;
    size_t count = 0;
    for (const auto& kv : byTime) {
        const size_t idx = kv.second;
        report() << "[" << count++ << "] " << kv.first << " usec"
                  << "\tavg: " << avg[idx] << "\tstddev: " << std::sqrt(var[idx] / numberTrials) << "\t";
        print_args(pargs[idx], extra[idx]);
    }
}

}  // namespace AppUtil
}  // namespace gat
