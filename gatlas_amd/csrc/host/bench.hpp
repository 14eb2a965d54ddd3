// Journal (memo file), Bench (one trial) and the AppUtil benchmark loops,
// with the behaviour and text formats of GatlasBenchmark.{hpp,cpp} and
// GatlasAppUtil.{hpp,cpp}.
#pragma once

#include <functional>
#include <map>
#include <string>
#include <vector>

#include "params.hpp"

namespace gat {

class DeviceApp;

// Journal record: "<kernelName>_<p0>_<p1>_..._<pn>_\t<int>\n".  Negative
// values are run states, values >= 0 are trial times in microseconds (0 =
// the trial ran but failed its check).
class Journal {
public:
    enum RunState {
        MISSING = 0,
        BUILD_IN_PROGRESS = -1,
        BUILD_OK = -2,
        RUN_IN_PROGRESS = -3,
        RUN_OK = -4,
    };
    explicit Journal(const std::string& file) : file_(file) {}

    static std::string key(const KernelInterface& k, const std::vector<size_t>& params);
    bool loadMemo();
    // rewrite keeping bad keys (no time) and, unless deleteTimes, the RUN_OK
    // state and every time of good keys; returns the number of bad keys or -1
    int purgeMemo(bool deleteTimes = false);
    size_t memoGood() const { return times_.size(); }
    int memoRunState(const KernelInterface& k, const std::vector<size_t>& params) const;
    int memoTime(const KernelInterface& k, const std::vector<size_t>& params, size_t trial) const;
    bool takeMemo(const KernelInterface& k, const std::vector<size_t>& params, int value) const;
    // the same records under a key built by the caller (trials that have no
    // KernelInterface, e.g. the grouped kernel's tile shapes)
    int memoTime(const std::string& key, size_t trial) const;
    bool takeMemo(const std::string& key, int value) const;
    const std::string& file() const { return file_; }

private:
    std::string file_;
    std::map<std::string, int> state_;
    std::map<std::string, std::vector<size_t>> times_;
};

class Bench {
public:
    Bench(DeviceApp& app, KernelInterface& k, bool printStatus = true)
        : app_(app), k_(k), journal_(nullptr), print_(printStatus) {}
    Bench(DeviceApp& app, KernelInterface& k, Journal& j, bool printStatus = true)
        : app_(app), k_(k), journal_(&j), print_(printStatus) {}
    bool printStatus() const { return print_; }
    // GatlasBenchmark.cpp:178-289: returns elapsed microseconds of
    // numTrials chained launches plus the wait, 0 on any failure.
    size_t run(size_t numTrials, const std::vector<size_t>& args, bool busTransferToDevice,
               bool busTransferFromDevice, bool printDebug = false);
    // Optional hook called between the BUILD_OK and RUN_IN_PROGRESS journal
    // records (used by failure-injection tests of the crash/hang recovery).
    std::function<void(const char* stage)> stageHook;
    // Each of the numTrials launches of run() is repeated `n` times back to
    // back and the elapsed time divided by n (default 1: one launch per
    // trial), so kernels of a few tens of microseconds are timed over enough
    // work for the host clock and launch overhead not to decide the ranking.
    void setLaunchesPerTrial(size_t n) { launches_ = n ? n : 1; }

private:
    DeviceApp& app_;
    KernelInterface& k_;
    Journal* journal_;
    bool print_;
    size_t launches_ = 1;
};

// Report stream of the benchmark loops: standard output, as the command-line
// tools print it, unless the calling thread has silenced it.
std::ostream& report();
// Silences report() on the calling thread while it lives (when `on`).
class QuietReports {
public:
    explicit QuietReports(bool on);
    ~QuietReports();

private:
    bool prev_;
};

namespace AppUtil {

void benchInit(const std::vector<std::vector<size_t>>& pargs, std::vector<bool>& ok, std::vector<size_t>& time,
               std::vector<size_t>& flops, std::vector<double>& avg, std::vector<double>& var,
               std::vector<std::vector<size_t>>& extra);

// GatlasAppUtil.cpp:175-283 (journal replay when `journal` is given,
// GatlasAppUtil.cpp:86-173 otherwise).  Returns the number of good kernels.
size_t benchLoop(size_t trialNumber, KernelInterface& kernel, Bench& bench, Journal* journal,
                 const std::vector<std::vector<size_t>>& pargs, std::vector<bool>& ok, std::vector<size_t>& time,
                 std::vector<size_t>& flops, std::vector<double>& avg, std::vector<double>& var,
                 std::vector<std::vector<size_t>>& extra, bool toDevice, bool fromDevice, bool dummyRun = false,
                 bool printDebug = false);

// keep the topN fastest accumulated times (ties kept together)
void markBench(size_t topN, std::vector<bool>& ok, const std::vector<size_t>& time);
// keep the topN highest average GFLOPS (ties kept together)
void markBenchAvg(size_t topN, std::vector<bool>& ok, const std::vector<double>& avg);
// index of the nth best average GFLOPS (ties: the last index wins), -1 if none
int rankBench(size_t nthPlace, const std::vector<bool>& ok, const std::vector<double>& avg);
void printBench(size_t numberTrials, const std::vector<std::vector<size_t>>& pargs, const std::vector<bool>& ok,
                const std::vector<size_t>& time, const std::vector<double>& avg, const std::vector<double>& var,
                const std::vector<std::vector<size_t>>& extra);

}  // namespace AppUtil

}  // namespace gat
