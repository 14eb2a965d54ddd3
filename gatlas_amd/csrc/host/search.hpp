// Autotuning search: candidate enumeration (getParams) and the two-step
// "EM" coordinate search of benchMatmul.cpp / benchMatvec.cpp /
// benchSaxpy.cpp, plus enumerators for the matrix-core and native kernels.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "bench.hpp"
#include "kernels_impl.hpp"

namespace gat {

constexpr long FREE = -1;

// benchMatmul.cpp:217-304 (buffer / image kernels)
std::vector<std::vector<size_t>> getParamsMatmul(KernelBaseMatmul& k, size_t maxWorkGroupSize, size_t vl,
                                                 size_t maxBlockHeight, size_t maxGroupSize, bool gemm, size_t M,
                                                 size_t N, size_t K, bool tA, bool tB, long groupSize,
                                                 long blockHeight, long extraParam);
// matrix-core kernels: every instantiated (waves, blocks, depth, shape)
// point, optionally restricted to fixed tiles (step 1) or a fixed extra
// parameter (step 0).  `fixed` holds (gH, gW, bh, bw) or is empty.
std::vector<std::vector<size_t>> getParamsMatmulMfma(KernelBaseMatmul& k, bool gemm, size_t M, size_t N, size_t K,
                                                     bool tA, bool tB, const std::vector<size_t>& fixed,
                                                     long extraParam);
// benchMatvec.cpp:202-289
std::vector<std::vector<size_t>> getParamsMatvec(KernelBaseMatvec& k, size_t maxWorkGroupSize, size_t vl,
                                                 size_t maxBlockHeight, size_t maxGroupSize, bool gemv, size_t M,
                                                 size_t N, bool tA, long groupSize, long blockHeight,
                                                 long extraParam);
std::vector<std::vector<size_t>> getParamsMatvecNative(KernelBaseMatvec& k, bool gemv, size_t M, size_t N, bool tA,
                                                       long groupSize, long blockHeight, long extraParam);
// benchSaxpy.cpp:158-275
std::vector<std::vector<size_t>> getParamsSaxpy(KernelBaseSaxpy& k, size_t maxWorkGroupSize, size_t vl,
                                                size_t maxBlockHeight, size_t maxGroupSize, size_t M, size_t N,
                                                long groupHeight, long groupWidth, long blockHeight,
                                                long blockWidth, long extraParam);
std::vector<std::vector<size_t>> getParamsSaxpyNative(KernelBaseSaxpy& k, size_t M, size_t N, long groupHeight,
                                                      long blockHeight, long extraParam);

struct LoopOpts {
    size_t numberTrials = 1;
    long topN = -1;
    bool toDevice = false, fromDevice = false, printDebug = false, paranoid = false;
};

// benchMatmul.cpp:307-405 mainLoop: trials x benchLoop, top-N pruning, ranking
// print.  Fills ok/avg; returns the good-kernel count of the last trial.
size_t mainLoop(KernelInterface& k, Bench& bench, Journal& journal, const std::vector<std::vector<size_t>>& pargs,
                std::vector<bool>& ok, std::vector<double>& avg, const LoopOpts& o);

// Two-step coordinate search (benchMatmul.cpp:536-627).  `enumerate(step,
// best)` returns the candidates of a step: step 0 frees the blocking with the
// extra parameter (and, for SAXPY, the blocking) of `best` fixed; step 1
// fixes the blocking of `best` and frees the rest.  `bumpExtra(best)` picks
// the next extra parameter when step 0 finds no good kernel (returns false
// to give up).  Returns the converged best parameter vector (empty if none).
std::vector<size_t> emSearch(KernelInterface& k, Bench& bench, Journal& journal, const LoopOpts& o,
                             const std::vector<size_t>& keyIndex,
                             const std::function<std::vector<std::vector<size_t>>(int, const std::vector<size_t>&)>&
                                 enumerate,
                             const std::function<bool(std::vector<size_t>&)>& bumpExtra);

// Search problem shared by the command-line tools and the C ABI.
struct SearchSpec {
    size_t maxWorkGroupSize = 1024;
    size_t vl = 1, maxBlockHeight = 10, maxGroupSize = 10;
    bool generalized = false, tA = false, tB = false;
    size_t M = 0, N = 0, K = 0;
};

// Full EM search of one kernel (any family/variant) for `spec`, followed by
// the final multi-trial average when o.numberTrials > 1.  Candidates whose
// journal record is complete are replayed without touching the device.
std::vector<size_t> emSearchKernel(KernelInterface& k, Bench& bench, Journal& journal, const LoopOpts& o,
                                   const SearchSpec& spec);
// Brute-force enumeration of all candidates (no -e), same conventions.
std::vector<std::vector<size_t>> allCandidates(KernelInterface& k, const SearchSpec& spec);

}  // namespace gat
