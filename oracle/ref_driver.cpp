// ref_driver.cpp -- golden-vector driver linked against the reference's own
// sources (compiled in place from the reference checkout by
// oracle/Makefile.ref; nothing here is copied from them).
//
// TEST INFRASTRUCTURE ONLY: runs in the build container to write fixtures
// under tests/golden/; never linked into or called by the product.
//
// It instantiates the reference's parameter classes (KernelBaseMatmul.cpp,
// KernelBaseMatvec.cpp, KernelBaseSaxpy.cpp), its Journal
// (GatlasBenchmark.cpp) and its ranking helpers (GatlasAppUtil.cpp), and
// answers line commands on stdin with one JSON object per line:
//   matmul <buffer|image> <p0..p11>          parameter functions
//   matvec <buffer|image> <p0..p8>
//   saxpy  <buffer|image> <p0..p8>
//   jfile <path> | jtake <name> <n> <p..> <v> | jload | jstate <name> <n> <p..>
//   jtime <name> <n> <p..> <trial> | jpurge <0|1> | jgood
//   rank <nth> <n> <ok..> <avg..> | mark <topN> <n> <ok..> <time..>
//   markavg <topN> <n> <ok..> <avg..>
// The reference kernel templates need an OpenCL device for setArgs, so the
// classes below register the same extra-parameter observers in the same
// order as KernelMatmulBuffer.hpp:42-49 / KernelMatmulImage.hpp:47-51 etc.
// and leave the device methods empty.
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "GatlasAppUtil.hpp"
#include "GatlasBenchmark.hpp"
#include "KernelBaseMatmul.hpp"
#include "KernelBaseMatvec.hpp"
#include "KernelBaseSaxpy.hpp"

using namespace gatlas;

#define NO_DEVICE                                                              \
    bool setArgs(OCLApp&, const size_t, const bool) { return false; }          \
    bool syncOutput(OCLApp&) { return false; }                                 \
    bool checkOutput(OCLApp&, const bool) { return false; }                    \
    void paranoidCheck() {}                                                    \
    std::ostream& print(std::ostream& os) const { return os; }                 \
    std::string kernelName() const { return name_; }                           \
    std::string name_;

struct MMBuf : KernelBaseMatmul, MatmulParamInlineMNK, MatmulParamLoopOrder {
    MMBuf() : MatmulParamInlineMNK(getExtraParameter()), MatmulParamLoopOrder(getExtraParameter()) {}
    NO_DEVICE
};
struct MMImg : KernelBaseMatmul, MatmulParamInlineMNK, MatmulParamLoopOrder, MatmulParamGlobalID {
    MMImg()
        : MatmulParamInlineMNK(getExtraParameter()),
          MatmulParamLoopOrder(getExtraParameter()),
          MatmulParamGlobalID(getExtraParameter()) {}
    NO_DEVICE
};
struct MVBuf : KernelBaseMatvec, MatvecParamInlineMN {
    MVBuf() : MatvecParamInlineMN(getExtraParameter()) {}
    NO_DEVICE
};
struct MVImg : KernelBaseMatvec, MatvecParamInlineMN, MatvecParamGlobalID {
    MVImg() : MatvecParamInlineMN(getExtraParameter()), MatvecParamGlobalID(getExtraParameter()) {}
    NO_DEVICE
};
struct SXBuf : KernelBaseSaxpy, SaxpyParamInlineMN {
    SXBuf() : SaxpyParamInlineMN(getExtraParameter()) {}
    NO_DEVICE
};
struct SXImg : KernelBaseSaxpy, SaxpyParamInlineMN, SaxpyParamGlobalID {
    SXImg() : SaxpyParamInlineMN(getExtraParameter()), SaxpyParamGlobalID(getExtraParameter()) {}
    NO_DEVICE
};

static void jvec(std::ostream& os, const std::vector<size_t>& v) {
    os << "[";
    for (size_t i = 0; i < v.size(); ++i) os << (i ? "," : "") << v[i];
    os << "]";
}

template <typename K>
static void report(K& k, const std::vector<size_t>& p) {
    k.setParams(p);
    std::vector<size_t> got;
    const bool valid = k.getParams(got);
    std::ostringstream os;
    os << "{\"in\":";
    jvec(os, p);
    os << ",\"valid\":" << (valid ? 1 : 0) << ",\"params\":";
    jvec(os, valid ? got : std::vector<size_t>());
    os << ",\"global\":";
    jvec(os, k.globalWorkItems());
    os << ",\"local\":";
    jvec(os, k.localWorkItems());
    os << ",\"flops\":" << k.numberFlops() << ",\"extra\":";
    jvec(os, k.extraParamDetail());
    os << ",\"total\":" << k.totalVariations() << "}";
    std::cout << os.str() << std::endl;
}

static std::vector<size_t> readn(std::istream& in, size_t n) {
    std::vector<size_t> v(n);
    for (size_t i = 0; i < n; ++i) in >> v[i];
    return v;
}

int main() {
    std::string line;
    std::unique_ptr<Journal> journal;
    MMBuf namer;  // any KernelInterface with a settable name serves the journal
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        if (!(in >> cmd)) continue;
        if (cmd == "matmul" || cmd == "matvec" || cmd == "saxpy") {
            std::string var;
            in >> var;
            const size_t n = cmd == "matmul" ? 12 : 9;
            std::vector<size_t> p = readn(in, n);
            if (cmd == "matmul") {
                if (var == "image") { MMImg k; report(k, p); }
                else { MMBuf k; report(k, p); }
            } else if (cmd == "matvec") {
                if (var == "image") { MVImg k; report(k, p); }
                else { MVBuf k; report(k, p); }
            } else {
                if (var == "image") { SXImg k; report(k, p); }
                else { SXBuf k; report(k, p); }
            }
        } else if (cmd == "jfile") {
            std::string path;
            in >> path;
            journal.reset(new Journal(path));
            std::cout << "{\"ok\":1}" << std::endl;
        } else if (cmd == "jtake" || cmd == "jstate" || cmd == "jtime") {
            size_t n;
            in >> namer.name_ >> n;
            std::vector<size_t> p = readn(in, n);
            long v = 0;
            in >> v;
            if (cmd == "jtake") std::cout << "{\"ok\":" << journal->takeMemo(namer, p, (int)v) << "}" << std::endl;
            else if (cmd == "jstate") std::cout << "{\"state\":" << journal->memoRunState(namer, p) << "}" << std::endl;
            else std::cout << "{\"time\":" << journal->memoTime(namer, p, (size_t)v) << "}" << std::endl;
        } else if (cmd == "jload") {
            std::cout << "{\"ok\":" << journal->loadMemo() << "}" << std::endl;
        } else if (cmd == "jpurge") {
            int d;
            in >> d;
            std::cout << "{\"bad\":" << journal->purgeMemo(d != 0) << "}" << std::endl;
        } else if (cmd == "jgood") {
            std::cout << "{\"good\":" << journal->memoGood() << "}" << std::endl;
        } else if (cmd == "rank" || cmd == "mark" || cmd == "markavg") {
            size_t a, n;
            in >> a >> n;
            std::vector<bool> ok(n);
            for (size_t i = 0; i < n; ++i) {
                int b;
                in >> b;
                ok[i] = b != 0;
            }
            std::ostringstream os;
            if (cmd == "mark") {
                std::vector<size_t> t = readn(in, n);
                AppUtil::markBench(a, ok, t);
            } else {
                std::vector<double> avg(n);
                for (size_t i = 0; i < n; ++i) in >> avg[i];
                if (cmd == "rank") {
                    std::cout << "{\"rank\":" << AppUtil::rankBench(a, ok, avg) << "}" << std::endl;
                    continue;
                }
                AppUtil::markBench(a, ok, avg);
            }
            os << "{\"ok\":[";
            for (size_t i = 0; i < n; ++i) os << (i ? "," : "") << (ok[i] ? 1 : 0);
            os << "]}";
            std::cout << os.str() << std::endl;
        } else {
            std::cout << "{\"error\":\"unknown command\"}" << std::endl;
        }
    }
    return 0;
}
