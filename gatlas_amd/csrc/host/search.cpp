#include "search.hpp"

#include <cmath>
#include <iostream>

#include "../context.hpp"

namespace gat {

std::vector<std::vector<size_t>> getParamsMatmul(KernelBaseMatmul& k, size_t maxWG, size_t vl, size_t maxBH,
                                                 size_t maxGS, bool gemm, size_t M, size_t N, size_t K, bool tA,
                                                 bool tB, long gs, long bh, long xp) {
    std::vector<std::vector<size_t>> pargs;
    std::vector<size_t> a;
    k.setGeneralizedMatmul(gemm);
    k.setMatrixDimensions(M, N, K);
    k.setDataLayout(tA, tB);
    if (gs != FREE) k.setWorkGroup((size_t)gs);
    if (bh != FREE) k.setInnerBlocking((size_t)bh, vl);
    if (xp != FREE) k.setExtraParameter((size_t)xp);
    auto push = [&]() {
        if (k.getParams(a)) pargs.push_back(a);
    };
    if (gs != FREE && bh != FREE && xp != FREE) {
        push();
    } else if (gs != FREE && bh != FREE) {
        for (size_t x = 0; x < k.totalVariations(); ++x) {
            k.setExtraParameter(x);
            push();
        }
    } else if (gs != FREE) {
        for (size_t b = vl; b <= maxBH; ++b) {
            k.setInnerBlocking(b, vl);
            for (size_t x = 0; x < k.totalVariations(); ++x) {
                k.setExtraParameter(x);
                push();
            }
        }
    } else {
        const size_t largestPossible = (size_t)std::sqrt((double)maxWG);
        const size_t largest = maxGS < largestPossible ? maxGS : largestPossible;
        const size_t bmin = bh != FREE ? (size_t)bh : vl, bmax = bh != FREE ? (size_t)bh : maxBH;
        const size_t xmin = xp != FREE ? (size_t)xp : 0, xmax = xp != FREE ? (size_t)xp + 1 : k.totalVariations();
        // largest valid group size for the problem, then the 8x8 group
        for (size_t wg = largest; wg > 8; --wg) {
            k.setWorkGroup(wg);
            bool any = false;
            for (size_t b = bmin; b <= bmax; ++b) {
                k.setInnerBlocking(b, vl);
                for (size_t x = xmin; x < xmax; ++x) {
                    k.setExtraParameter(x);
                    if (k.getParams(a)) {
                        pargs.push_back(a);
                        any = true;
                    }
                }
            }
            if (any) break;
        }
        k.setWorkGroup(8);
        for (size_t b = bmin; b <= bmax; ++b) {
            k.setInnerBlocking(b, vl);
            for (size_t x = xmin; x < xmax; ++x) {
                k.setExtraParameter(x);
                push();
            }
        }
    }
    return pargs;
}

std::vector<std::vector<size_t>> getParamsMatmulMfma(KernelBaseMatmul& k, bool gemm, size_t M, size_t N, size_t K,
                                                     bool tA, bool tB, const std::vector<size_t>& fixed, long xp) {
    std::vector<std::vector<size_t>> pargs;
    std::vector<size_t> a;
    k.setGeneralizedMatmul(gemm);
    k.setMatrixDimensions(M, N, K);
    k.setDataLayout(tA, tB);
    const int dt = kernel_dtype(&k);
    const int n = Context::num_cfgs(dt);
    // candidate (waves, blocks) in table order, deduplicated
    std::vector<std::vector<size_t>> tiles;
    for (int i = 0; i < n; ++i) {
        const MfmaCfg c = Context::cfg_at(dt, i);
        std::vector<size_t> t{(size_t)c.waves_m, (size_t)c.waves_n, (size_t)c.tm, (size_t)c.tn};
        bool seen = false;
        for (auto& u : tiles) seen = seen || u == t;
        if (!seen) tiles.push_back(t);
    }
    for (const auto& t : tiles) {
        if (!fixed.empty() && t != fixed) continue;
        k.setWorkGroup(t[0], t[1]);
        k.setInnerBlocking(t[2], t[3]);
        const size_t xmin = xp != FREE ? (size_t)xp : 0, xmax = xp != FREE ? (size_t)xp + 1 : k.totalVariations();
        for (size_t x = xmin; x < xmax; ++x) {
            k.setExtraParameter(x);
            if (k.getParams(a)) pargs.push_back(a);
        }
    }
    return pargs;
}

std::vector<std::vector<size_t>> getParamsMatvec(KernelBaseMatvec& k, size_t maxWG, size_t vl, size_t maxBH,
                                                 size_t maxGS, bool gemv, size_t M, size_t N, bool tA, long gs,
                                                 long bh, long xp) {
    std::vector<std::vector<size_t>> pargs;
    std::vector<size_t> a;
    k.setGeneralizedMatvec(gemv);
    k.setMatrixDimensions(M, N);
    k.setDataLayout(tA);
    if (gs != FREE) k.setWorkGroup((size_t)gs);
    if (bh != FREE) k.setInnerBlocking((size_t)bh, vl);
    if (xp != FREE) k.setExtraParameter((size_t)xp);
    auto push = [&]() {
        if (k.getParams(a)) pargs.push_back(a);
    };
    if (gs != FREE && bh != FREE && xp != FREE) {
        push();
    } else if (gs != FREE && bh != FREE) {
        for (size_t x = 0; x < k.totalVariations(); ++x) {
            k.setExtraParameter(x);
            push();
        }
    } else if (gs != FREE) {
        for (size_t b = vl; b <= maxBH; ++b) {
            k.setInnerBlocking(b, vl);
            for (size_t x = 0; x < k.totalVariations(); ++x) {
                k.setExtraParameter(x);
                push();
            }
        }
    } else {
        const size_t largest = maxGS < maxWG ? maxGS : maxWG;
        const size_t bmin = bh != FREE ? (size_t)bh : vl, bmax = bh != FREE ? (size_t)bh : maxBH;
        const size_t xmin = xp != FREE ? (size_t)xp : 0, xmax = xp != FREE ? (size_t)xp + 1 : k.totalVariations();
        for (size_t wg = largest; wg > 64; --wg) {
            k.setWorkGroup(wg);
            bool any = false;
            for (size_t b = bmin; b <= bmax; ++b) {
                k.setInnerBlocking(b, vl);
                for (size_t x = xmin; x < xmax; ++x) {
                    k.setExtraParameter(x);
                    if (k.getParams(a)) {
                        pargs.push_back(a);
                        any = true;
                    }
                }
            }
            if (any) break;
        }
        k.setWorkGroup(64);
        for (size_t b = bmin; b <= bmax; ++b) {
            k.setInnerBlocking(b, vl);
            for (size_t x = xmin; x < xmax; ++x) {
                k.setExtraParameter(x);
                push();
            }
        }
    }
    return pargs;
}

std::vector<std::vector<size_t>> getParamsMatvecNative(KernelBaseMatvec& k, bool gemv, size_t M, size_t N, bool tA,
                                                       long gs, long bh, long xp) {
    std::vector<std::vector<size_t>> pargs;
    std::vector<size_t> a;
    k.setGeneralizedMatvec(gemv);
    k.setMatrixDimensions(M, N);
    k.setDataLayout(tA);
    const size_t vl = kernel_dtype(&k) == DT_F64 ? 2 : 4;
    static const size_t wgs[] = {512, 256, 128, 64};
    static const size_t rows[] = {1, 2, 4, 8};
    for (size_t w : wgs) {
        if (gs != FREE && (size_t)gs != w) continue;
        for (size_t r : rows) {
            if (bh != FREE && (size_t)bh != r) continue;
            k.setWorkGroup(w);
            k.setInnerBlocking(r, vl);
            const size_t xmin = xp != FREE ? (size_t)xp : 0, xmax = xp != FREE ? (size_t)xp + 1 : k.totalVariations();
            for (size_t x = xmin; x < xmax; ++x) {
                k.setExtraParameter(x);
                if (k.getParams(a)) pargs.push_back(a);
            }
        }
    }
    return pargs;
}

std::vector<std::vector<size_t>> getParamsSaxpy(KernelBaseSaxpy& k, size_t maxWG, size_t vl, size_t maxBH,
                                                size_t maxGS, size_t M, size_t N, long gH, long gW, long bh, long bw,
                                                long xp) {
    std::vector<std::vector<size_t>> pargs;
    std::vector<size_t> a;
    k.setSaxpyDimensions(M, N);
    k.setVectorLength(vl);
    if (gH != FREE) k.setWorkGroup((size_t)gH, (size_t)gW);
    if (bh != FREE) k.setInnerBlocking((size_t)bh, (size_t)bw);
    if (xp != FREE) k.setExtraParameter((size_t)xp);
    auto push = [&]() {
        if (k.getParams(a)) pargs.push_back(a);
    };
    const size_t largest = maxGS < maxWG ? maxGS : maxWG;
    if (gH != FREE && bh != FREE && xp != FREE) {
        push();
    } else if (gH != FREE && bh == FREE && xp == FREE) {
        if (vl == N) {
            for (size_t b = 1; b <= maxBH; ++b) {
                k.setInnerBlocking(b, 0);
                for (size_t x = 0; x < k.totalVariations(); ++x) {
                    k.setExtraParameter(x);
                    push();
                }
            }
        } else {
            for (size_t b = 1; b <= maxBH; ++b)
                for (size_t w = 1; w <= 8; ++w) {
                    k.setInnerBlocking(b, w);
                    for (size_t x = 0; x < k.totalVariations(); ++x) {
                        k.setExtraParameter(x);
                        push();
                    }
                }
        }
    } else if (gH == FREE && bh != FREE && xp != FREE) {
        if (vl == N) {
            for (size_t wg = 64; wg <= largest; ++wg) {
                k.setWorkGroup(wg, 0);
                push();
            }
        } else {
            for (size_t h = 1; h <= largest; ++h)
                for (size_t w = 1; w <= largest; ++w)
                    if (h * w <= largest) {
                        k.setWorkGroup(h, w);
                        push();
                    }
        }
    } else {
        const size_t bmin = bh != FREE ? (size_t)bh : 1, bmax = bh != FREE ? (size_t)bh : maxBH;
        const size_t xmin = xp != FREE ? (size_t)xp : 0, xmax = xp != FREE ? (size_t)xp + 1 : k.totalVariations();
        if (vl == N) {
            for (size_t wg = 64; wg <= largest; ++wg) {
                k.setWorkGroup(wg, 0);
                for (size_t b = bmin; b <= bmax; ++b) {
                    k.setInnerBlocking(b, 0);
                    for (size_t x = xmin; x < xmax; ++x) {
                        k.setExtraParameter(x);
                        push();
                    }
                }
            }
        } else {
            for (size_t h = 1; h <= largest; ++h)
                for (size_t w = 1; w <= largest; ++w)
                    if (h * w <= largest) {
                        k.setWorkGroup(h, w);
                        for (size_t b = bmin; b <= bmax; ++b)
                            for (size_t bwv = 1; bwv <= 8; ++bwv) {
                                k.setInnerBlocking(b, bwv);
                                for (size_t x = xmin; x < xmax; ++x) {
                                    k.setExtraParameter(x);
                                    push();
                                }
                            }
                    }
        }
    }
    return pargs;
}

std::vector<std::vector<size_t>> getParamsSaxpyNative(KernelBaseSaxpy& k, size_t M, size_t N, long gH, long bh,
                                                      long xp) {
    std::vector<std::vector<size_t>> pargs;
    std::vector<size_t> a;
    const size_t vl = kernel_dtype(&k) == DT_F64 ? 2 : 4;
    k.setSaxpyDimensions(M, N);
    k.setVectorLength(vl);
    static const size_t wgs[] = {64, 128, 256, 512, 1024};
    static const size_t un[] = {1, 2, 4, 8};
    for (size_t w : wgs) {
        if (gH != FREE && (size_t)gH != w) continue;
        for (size_t u : un) {
            if (bh != FREE && (size_t)bh != u) continue;
            k.setWorkGroup(w, 0);
            k.setInnerBlocking(u, 0);
            const size_t xmin = xp != FREE ? (size_t)xp : 0, xmax = xp != FREE ? (size_t)xp + 1 : k.totalVariations();
            for (size_t x = xmin; x < xmax; ++x) {
                k.setExtraParameter(x);
                if (k.getParams(a)) pargs.push_back(a);
            }
        }
    }
    return pargs;
}

size_t mainLoop(KernelInterface& k, Bench& bench, Journal& journal, const std::vector<std::vector<size_t>>& pargs,
                std::vector<bool>& ok, std::vector<double>& avg, const LoopOpts& o) {
    std::vector<size_t> time, flops;
    std::vector<double> var;
    std::vector<std::vector<size_t>> extra;
    std::vector<bool> ok0;
    std::vector<double> avg0;
    AppUtil::benchInit(pargs, ok0, time, flops, avg0, var, extra);
    if (ok.size() != pargs.size()) ok.assign(pargs.size(), true);
    if (avg.size() != pargs.size()) avg.assign(pargs.size(), 0.0);
    if (o.paranoid) k.paranoidCheck();
    size_t good = 0;
    for (size_t t = 0; t < o.numberTrials; ++t) {
        good = AppUtil::benchLoop(t, k, bench, &journal, pargs, ok, time, flops, avg, var, extra, o.toDevice,
                                  o.fromDevice, t == 0, o.printDebug);
        if (o.topN != -1) AppUtil::markBench((size_t)o.topN, ok, time);
    }
    AppUtil::printBench(o.numberTrials, pargs, ok, time, avg, var, extra);
    return good;
}

std::vector<size_t> emSearch(KernelInterface& k, Bench& bench, Journal& journal, const LoopOpts& o,
                             const std::vector<size_t>& keyIndex,
                             const std::function<std::vector<std::vector<size_t>>(int, const std::vector<size_t>&)>&
                                 enumerate,
                             const std::function<bool(std::vector<size_t>&)>& bumpExtra) {
    LoopOpts one = o;
    one.numberTrials = 1;
    std::vector<size_t> best;  // the current best parameter vector (empty before the first)
    bool found = false;
    int guard = 0;
    while (!found) {
        for (int step = 0; step <= 1; ++step) {
            int bestIndex = -1;
            while (bestIndex == -1) {
                std::vector<std::vector<size_t>> pargs = enumerate(step, best);
                std::vector<bool> ok(pargs.size(), true);
                std::vector<double> avg(pargs.size(), 0.0);
                journal.loadMemo();
                mainLoop(k, bench, journal, pargs, ok, avg, one);
                bestIndex = AppUtil::rankBench(0, ok, avg);
                if (bestIndex == -1) {
                    if (step == 0 && bumpExtra && bumpExtra(best) && ++guard < 4096) continue;
                    std::cerr << "error: no good kernels found so giving up" << std::endl;
                    return std::vector<size_t>();
                }
                guard = 0;
                const std::vector<size_t>& p = pargs[bestIndex];
                bool same = !best.empty();
                for (size_t i : keyIndex) same = same && best[i] == p[i];
                if (same) found = true;
                best = p;
                k.setParams(best);
            }
        }
    }
    return best;
}

std::vector<std::vector<size_t>> allCandidates(KernelInterface& k, const SearchSpec& s) {
    const int var = kernel_variant(&k);
    if (auto* m = as_matmul(&k)) {
        if (var == VAR_MFMA)
            return getParamsMatmulMfma(*m, s.generalized, s.M, s.N, s.K, s.tA, s.tB, std::vector<size_t>(), FREE);
        return getParamsMatmul(*m, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.generalized, s.M,
                               s.N, s.K, s.tA, s.tB, FREE, FREE, FREE);
    }
    if (auto* v = as_matvec(&k)) {
        if (var == VAR_NATIVE) return getParamsMatvecNative(*v, s.generalized, s.M, s.N, s.tA, FREE, FREE, FREE);
        return getParamsMatvec(*v, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.generalized, s.M,
                               s.N, s.tA, FREE, FREE, FREE);
    }
    if (auto* x = as_saxpy(&k)) {
        if (var == VAR_NATIVE) return getParamsSaxpyNative(*x, s.M, s.N, FREE, FREE, FREE);
        return getParamsSaxpy(*x, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.M, s.N, FREE, FREE,
                              FREE, FREE, FREE);
    }
    return {};
}

std::vector<size_t> emSearchKernel(KernelInterface& k, Bench& bench, Journal& journal, const LoopOpts& o,
                                   const SearchSpec& s) {
    const int var = kernel_variant(&k);
    std::vector<size_t> best;
    if (auto* m = as_matmul(&k)) {
        // benchMatmul.cpp:536-627: step 0 fixes the extra parameter, step 1
        // fixes the work group and blocking.
        const size_t total = m->totalVariations();
        size_t seedXp = 0;
        auto enumerate = [&](int step, const std::vector<size_t>& b) {
            if (var == VAR_MFMA) {
                if (step == 0)
                    return getParamsMatmulMfma(*m, s.generalized, s.M, s.N, s.K, s.tA, s.tB, std::vector<size_t>(),
                                               (long)(b.empty() ? seedXp : b[11]));
                return getParamsMatmulMfma(*m, s.generalized, s.M, s.N, s.K, s.tA, s.tB,
                                           std::vector<size_t>{b[7], b[8], b[9], b[10]}, FREE);
            }
            if (step == 0)
                return getParamsMatmul(*m, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.generalized,
                                       s.M, s.N, s.K, s.tA, s.tB, FREE, FREE, (long)(b.empty() ? seedXp : b[11]));
            return getParamsMatmul(*m, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.generalized,
                                   s.M, s.N, s.K, s.tA, s.tB, (long)b[7], (long)b[9], FREE);
        };
        auto bump = [&](std::vector<size_t>& b) {
            if (!b.empty()) {
                b[11] = (b[11] + 1) % total;
                return true;
            }
            return ++seedXp < total;
        };
        best = emSearch(k, bench, journal, o, {7, 8, 9, 10, 11}, enumerate, bump);
    } else if (auto* v = as_matvec(&k)) {
        // benchMatvec.cpp:521-612
        const size_t total = v->totalVariations();
        size_t seedXp = 0;
        auto params = [&](long g, long b, long x) {
            if (var == VAR_NATIVE) return getParamsMatvecNative(*v, s.generalized, s.M, s.N, s.tA, g, b, x);
            return getParamsMatvec(*v, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.generalized, s.M,
                                   s.N, s.tA, g, b, x);
        };
        auto enumerate = [&](int step, const std::vector<size_t>& b) {
            if (step == 0) return params(FREE, FREE, (long)(b.empty() ? seedXp : b[8]));
            return params((long)b[5], (long)b[6], FREE);
        };
        auto bump = [&](std::vector<size_t>& b) {
            if (!b.empty()) {
                b[8] = (b[8] + 1) % total;
                return true;
            }
            return ++seedXp < total;
        };
        best = emSearch(k, bench, journal, o, {5, 6, 8}, enumerate, bump);
    } else if (auto* x = as_saxpy(&k)) {
        // benchSaxpy.cpp:498-584: step 0 frees the work group with blocking
        // and extra parameter fixed; step 1 frees blocking and extra parameter.
        auto params = [&](long gH, long gW, long bh, long bw, long xp) {
            if (var == VAR_NATIVE) return getParamsSaxpyNative(*x, s.M, s.N, gH, bh, xp);
            return getParamsSaxpy(*x, s.maxWorkGroupSize, s.vl, s.maxBlockHeight, s.maxGroupSize, s.M, s.N, gH, gW,
                                  bh, bw, xp);
        };
        const long initBH = var == VAR_NATIVE ? 4 : 1;
        const long initBW = (s.N == s.vl) ? 0 : (long)s.vl;
        auto enumerate = [&](int step, const std::vector<size_t>& b) {
            if (step == 0) {
                if (b.empty()) return params(FREE, FREE, initBH, initBW, 0);
                return params(FREE, FREE, (long)b[5], (long)b[6], (long)b[8]);
            }
            return params((long)b[3], (long)b[4], FREE, FREE, FREE);
        };
        best = emSearch(k, bench, journal, o, {3, 4, 5, 6, 8}, enumerate, nullptr);
    }
    if (!best.empty() && o.numberTrials > 1) {
        std::vector<std::vector<size_t>> pargs{best};
        std::vector<bool> ok(1, true);
        std::vector<double> avg(1, 0.0);
        journal.loadMemo();
        mainLoop(k, bench, journal, pargs, ok, avg, o);
    }
    return best;
}

}  // namespace gat
