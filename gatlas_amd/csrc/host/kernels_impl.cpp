#include "kernels_impl.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <sstream>

#include "../kernels.hpp"
#include "checks.hpp"

namespace gat {

const int kMfmaBK[4] = {8, 16, 32, 64};
// FP8 tiles are always 128 deep (one scale block), so their depth digit has
// the single value 128; the other forms take kMfmaBK.
size_t mfma_depths(int dtype) { return dtype == DT_FP8 ? 1 : 4; }
int mfma_depth(int dtype, size_t i) { return dtype == DT_FP8 ? 128 : kMfmaBK[i]; }
const int kMfmaBlk[2] = {32, 16};
const int kMfmaWK[3] = {1, 2, 4};

template <typename T>
static const char* nameof();
template <>
const char* nameof<float>() { return "float"; }
template <>
const char* nameof<double>() { return "double"; }
template <typename T>
static int dtypeof();
template <>
int dtypeof<float>() { return DT_F32; }
template <>
int dtypeof<double>() { return DT_F64; }

static void print_vec(std::ostream& os, const std::vector<size_t>& v) {
    os << "(";
    for (size_t i = 0; i < v.size(); ++i) os << (i ? ", " : "") << v[i];
    os << ")";
}

// Tag interface for the factory's typed accessors.
struct KernelTag {
    int family, variant, dtype;
    virtual ~KernelTag() {}
};

// ======================================================================
// matmul, reference geometry (buffer / image)
// ======================================================================
template <typename T>
class KMatmulCompat : public KernelBaseMatmul, public KernelTag {
    int vl_;
    int hA_ = -1, hB_ = -1, hC_ = -1;
    bool paranoid_ = false;
    std::vector<T> ref_;
    T alpha_ = 1, beta_ = 1;
    double lastErr_ = 0;

public:
    KMatmulCompat(int variant, int vl) : vl_(vl) {
        family = FAM_MATMUL;
        this->variant = variant;
        dtype = dtypeof<T>();
        xp_.addDigit(2);  // inlineMNK
        xp_.addDigit(6);  // loopOrder
        if (variant == VAR_IMAGE) xp_.addDigit(2);  // globalID
    }
    std::string kernelName() const override {
        std::ostringstream s;
        s << (variant == VAR_IMAGE ? "matmulimage" : "matmulbuffer") << nameof<T>() << vl_;
        return s.str();
    }
    void paranoidCheck() override { paranoid_ = true; }
    double lastCheckError() const override { return lastErr_; }

    bool setArgs(DeviceApp& app, bool syncInput) override {
        const size_t P = packedCalc(), M = dimM(), N = dimN(), K = dimK();
        if (hA_ < 0 || hB_ < 0 || hC_ < 0 || dimChanged() || packedChanged() ||
            app.bytes(hC_) != P * M * N * sizeof(T)) {
            app.releaseBuffers();
            hA_ = app.createBuffer<T>(P * M * K, "matA", (T)1);
            hB_ = app.createBuffer<T>(P * K * N, "matB", (T)1);
            hC_ = app.createBuffer<T>(P * M * N, "matC", (T)0);
            if (hA_ < 0 || hB_ < 0 || hC_ < 0) return false;
        } else {
            if (syncInput && (!app.syncToDevice(hA_) || !app.syncToDevice(hB_))) return false;
            if (!app.clearBuffer(hC_)) return false;
        }
        alpha_ = (paranoid_ && generalizedMatmul()) ? checks::posrand<T>() : (T)1;
        beta_ = (paranoid_ && generalizedMatmul()) ? checks::posrand<T>() : (T)1;
        if (paranoid_) {
            T* A = app.bufferPtr<T>(hA_);
            T* B = app.bufferPtr<T>(hB_);
            T* C = app.bufferPtr<T>(hC_);
            checks::fillrand(A, P * M * K);
            if (!app.syncToDevice(hA_)) return false;
            checks::fillrand(B, P * K * N);
            if (!app.syncToDevice(hB_)) return false;
            if (generalizedMatmul()) {
                checks::fillrand(C, P * M * N);
                if (!app.syncToDevice(hC_)) return false;
            }
            ref_.assign(P * M * N, (T)0);
            checks::refMatmul<T>(P, generalizedMatmul(), transposeA(), transposeB(), M, N, K, alpha_, beta_, A, B,
                                 generalizedMatmul() ? C : nullptr, ref_.data());
        }
        return true;
    }
    bool build(DeviceApp&) override {
        return groupSize() > 0 && blockWidth() == (size_t)vl_ && validParams() &&
               compat_has(dtype, CompatCfg{(int)groupSize(), (int)blockHeight(), vl_});
    }
    bool enqueue(DeviceApp& app) override {
        GemmArgs a;
        a.transA = transposeA();
        a.transB = transposeB();
        a.M = dimM();
        a.N = dimN();
        a.K = dimK();
        a.A = app.devicePtr(hA_);
        a.B = app.devicePtr(hB_);
        a.C = app.devicePtr(hC_);
        a.lda = transposeA() ? a.M : a.K;
        a.ldb = transposeB() ? a.K : a.N;
        a.ldc = a.N;
        a.strideA = a.M * a.K;
        a.strideB = a.K * a.N;
        a.strideC = a.M * a.N;
        a.batch = (int)packedCalc();
        a.alpha = generalizedMatmul() ? (double)alpha_ : 1.0;
        a.beta = generalizedMatmul() ? (double)beta_ : 0.0;
        return launch_gemm_compat(dtype, CompatCfg{(int)groupSize(), (int)blockHeight(), vl_}, a, app.stream()) ==
               ST_OK;
    }
    bool syncOutput(DeviceApp& app) override { return app.syncFromDevice(hC_); }
    bool checkOutput(DeviceApp& app, bool printOutput) override {
        const size_t n = packedCalc() * dimM() * dimN();
        const T* C = app.bufferPtr<T>(hC_);
        if (printOutput)
            for (size_t i = 0; i < std::min<size_t>(n, 64); ++i) std::cerr << C[i] << (i + 1 < n ? " " : "\n");
        if (paranoid_) {
            long double d = 0;
            const bool ok = checks::paranoidPass(C, ref_.data(), n, &d);
            lastErr_ = (double)d;
            std::cerr << "absdiff: " << (double)d << "\t";
            return ok;
        }
        return checks::checkArray(C, n, (T)dimK());
    }
    std::ostream& print(std::ostream& os) const override {
        os << "// " << kernelName() << ": gemm_compat_kernel<" << nameof<T>() << ", VL=" << vl_
           << ", BH=" << blockHeight() << ">\n"
           << "//   work-group " << groupHeight() << " x " << groupWidth() << ", macro tile "
           << groupHeight() * blockHeight() << " x " << groupWidth() * blockWidth() << ", K step "
           << groupHeight() * vl_ << "\n//   global work-items ";
        print_vec(os, globalWorkItems());
        os << ", local ";
        print_vec(os, localWorkItems());
        os << "\n//   M=" << dimM() << " N=" << dimN() << " K=" << dimK() << " transA=" << transposeA()
           << " transB=" << transposeB() << " GEMM=" << generalizedMatmul() << " packed=" << packedCalc()
           << "\n//   per element: k-ascending chain acc = acc + a*b (separately rounded), epilogue "
           << (generalizedMatmul() ? "alpha*acc + beta*C" : "C = acc") << "\n";
        return os;
    }
};

// ======================================================================
// matmul, matrix cores
// ======================================================================
class KMatmulMfma : public KernelBaseMatmul, public KernelTag {
    int hA_ = -1, hB_ = -1, hC_ = -1;
    int hA2_ = -1, hB2_ = -1, hSA_ = -1, hSB_ = -1;  // converted operands (BF16 / FP8)
    bool paranoid_ = false;
    bool f64_;
    double alpha_ = 1, beta_ = 1;
    std::vector<size_t> rows_;
    std::vector<double> ref_, mag_;
    double lastErr_ = 0;

    template <typename T>
    void prepare(DeviceApp& app) {
        const size_t P = packedCalc(), M = dimM(), N = dimN(), K = dimK();
        T* A = app.bufferPtr<T>(hA_);
        T* B = app.bufferPtr<T>(hB_);
        T* C = app.bufferPtr<T>(hC_);
        checks::fillrand(A, P * M * K);
        checks::fillrand(B, P * K * N);
        if (generalizedMatmul()) checks::fillrand(C, P * M * N);
        // FP64 reference on a row sample of every packed problem
        rows_.clear();
        const size_t ns = std::min<size_t>(M, 24);
        for (size_t s = 0; s < ns; ++s) rows_.push_back(ns > 1 ? s * (M - 1) / (ns - 1) : 0);
        rows_.erase(std::unique(rows_.begin(), rows_.end()), rows_.end());
        ref_.assign(P * rows_.size() * N, 0.0);
        mag_.assign(P * rows_.size() * N, 0.0);
        std::vector<double> a64(M * K), b64(K * N), c64(M * N);
        for (size_t p = 0; p < P; ++p) {
            for (size_t i = 0; i < M * K; ++i) a64[i] = A[p * M * K + i];
            for (size_t i = 0; i < K * N; ++i) b64[i] = B[p * K * N + i];
            for (size_t i = 0; i < M * N; ++i) c64[i] = generalizedMatmul() ? (double)C[p * M * N + i] : 0.0;
            for (size_t r = 0; r < rows_.size(); ++r) {
                const size_t off = (p * rows_.size() + r) * N;
                checks::gemm64Rows(transposeA(), transposeB(), M, N, K, alpha_, a64.data(), b64.data(),
                                   generalizedMatmul() ? beta_ : 0.0, c64.data(), rows_[r], rows_[r] + 1,
                                   &ref_[off], &mag_[off]);
            }
        }
    }
    template <typename T>
    bool check(DeviceApp& app) {
        const size_t P = packedCalc(), M = dimM(), N = dimN();
        const T* C = app.bufferPtr<T>(hC_);
        if (!paranoid_) return checks::checkArray(C, P * M * N, (T)dimK());
        std::vector<double> got(ref_.size());
        for (size_t p = 0; p < P; ++p)
            for (size_t r = 0; r < rows_.size(); ++r)
                for (size_t j = 0; j < N; ++j)
                    got[(p * rows_.size() + r) * N + j] = C[p * M * N + rows_[r] * N + j];
        lastErr_ = checks::componentwiseError(got.data(), ref_.data(), mag_.data(), got.size());
        std::cerr << "componentwise err: " << lastErr_ << "\t";
        return lastErr_ <= tolerance();
    }

public:
    explicit KMatmulMfma(int dt) : f64_(dt == DT_F64) {
        family = FAM_MATMUL;
        variant = VAR_MFMA;
        dtype = dt;
        xp_.addDigit(mfma_depths(dt));  // LDS depth bk (mfma_depth)
        xp_.addDigit(2);  // block shape in kMfmaBlk
        xp_.addDigit(3);  // wave groups sharing each K tile in kMfmaWK
    }
    double tolerance() const {
        switch (dtype) {
            case DT_F64: return 1e-12;
            case DT_F32: return 1e-5;
            case DT_BF16X3: return 1e-5;
            case DT_BF16: return 2e-2;
            case DT_FP8: return 2.5e-1;
        }
        return 0.0;
    }
    std::string kernelName() const override {
        static const char* nm[] = {"float", "double", "bf16", "bf16x3", "fp8"};
        return std::string("matmulmfma") + nm[dtype];
    }
    MfmaCfg cfg() const {
        const auto d = extraParamDetail();
        return MfmaCfg{(int)groupHeight(), (int)groupWidth(), (int)blockHeight(), (int)blockWidth(),
                       mfma_depth(dtype, d.empty() ? 0 : d[0]), kMfmaBlk[d.size() < 2 ? 0 : d[1]],
                       kMfmaWK[d.size() < 3 ? 0 : d[2]]};
    }
    bool validParams() const override {
        return packedCalc() > 0 && dimM() > 0 && dimN() > 0 && dimK() > 0 && extraParam() < totalVariations() &&
               Context::has_cfg(dtype, cfg());
    }
    std::vector<size_t> globalWorkItems() const override {
        const MfmaCfg c = cfg();
        const size_t BM = (size_t)c.waves_m * c.tm * c.blk, BN = (size_t)c.waves_n * c.tn * c.blk;
        if (!BM || !BN) return {0, 0};
        const size_t tiles = ((dimM() + BM - 1) / BM) * ((dimN() + BN - 1) / BN);
        return {tiles * (size_t)c.waves_m * c.waves_n * c.wk * 64, packedCalc()};
    }
    std::vector<size_t> localWorkItems() const override {
        return {groupHeight() * groupWidth() * (size_t)cfg().wk * 64, 1};
    }
    void paranoidCheck() override { paranoid_ = true; }
    double lastCheckError() const override { return lastErr_; }

    bool setArgs(DeviceApp& app, bool syncInput) override {
        const size_t P = packedCalc(), M = dimM(), N = dimN(), K = dimK();
        const size_t es = f64_ ? 8 : 4;
        if (hA_ < 0 || hB_ < 0 || hC_ < 0 || dimChanged() || packedChanged() || app.bytes(hC_) != P * M * N * es) {
            app.releaseBuffers();
            hA2_ = hB2_ = hSA_ = hSB_ = -1;
            if (f64_) {
                hA_ = app.createBuffer<double>(P * M * K, "matA", 1.0);
                hB_ = app.createBuffer<double>(P * K * N, "matB", 1.0);
                hC_ = app.createBuffer<double>(P * M * N, "matC", 0.0);
            } else {
                hA_ = app.createBuffer<float>(P * M * K, "matA", 1.0f);
                hB_ = app.createBuffer<float>(P * K * N, "matB", 1.0f);
                hC_ = app.createBuffer<float>(P * M * N, "matC", 0.0f);
            }
            if (hA_ < 0 || hB_ < 0 || hC_ < 0) return false;
        } else {
            if (syncInput && (!app.syncToDevice(hA_) || !app.syncToDevice(hB_))) return false;
            if (!app.clearBuffer(hC_)) return false;
        }
        alpha_ = (paranoid_ && generalizedMatmul()) ? checks::posrand<double>() : 1.0;
        beta_ = (paranoid_ && generalizedMatmul()) ? checks::posrand<double>() : 1.0;
        if (!f64_) {
            alpha_ = (double)(float)alpha_;
            beta_ = (double)(float)beta_;
        }
        if (paranoid_) {
            if (f64_) prepare<double>(app);
            else prepare<float>(app);
            if (!app.syncToDevice(hA_) || !app.syncToDevice(hB_) || !app.syncToDevice(hC_)) return false;
        }
        return convertInputs(app);
    }
    // BF16: device bf16 copies of A and B; FP8: block-quantised op(A) (M x K)
    // and op(B)^T (N x K) with their scales.  Done here, outside the timed
    // launch, as the reference uploads its operands in setArgs.
    bool convertInputs(DeviceApp& app) {
        if (dtype != DT_BF16 && dtype != DT_FP8) return true;
        const size_t P = packedCalc(), M = dimM(), N = dimN(), K = dimK();
        if (dtype == DT_BF16) {
            if (hA2_ < 0) hA2_ = app.createRaw(P * M * K * 2, "matA.bf16");
            if (hB2_ < 0) hB2_ = app.createRaw(P * K * N * 2, "matB.bf16");
            if (hA2_ < 0 || hB2_ < 0) return false;
            return launch_convert_f32_bf16((const float*)app.devicePtr(hA_), app.devicePtr(hA2_), P * M * K,
                                           app.stream()) == ST_OK &&
                   launch_convert_f32_bf16((const float*)app.devicePtr(hB_), app.devicePtr(hB2_), P * K * N,
                                           app.stream()) == ST_OK &&
                   app.wait();
        }
        if (P != 1) return false;
        const size_t kb = (K + 127) / 128;
        if (hA2_ < 0) hA2_ = app.createRaw(M * K, "matA.fp8");
        if (hB2_ < 0) hB2_ = app.createRaw(N * K, "matB.fp8");
        if (hSA_ < 0) hSA_ = app.createRaw(((M + 127) / 128) * kb * 4, "scaleA");
        if (hSB_ < 0) hSB_ = app.createRaw(((N + 127) / 128) * kb * 4, "scaleB");
        if (hA2_ < 0 || hB2_ < 0 || hSA_ < 0 || hSB_ < 0) return false;
        const int64_t lda = transposeA() ? M : K, ldb = transposeB() ? K : N;
        return launch_quantize_fp8((const float*)app.devicePtr(hA_), transposeA() ? 1 : lda, transposeA() ? lda : 1,
                                   (int)M, (int)K, app.devicePtr(hA2_), (float*)app.devicePtr(hSA_),
                                   app.stream()) == ST_OK &&
               launch_quantize_fp8((const float*)app.devicePtr(hB_), transposeB() ? ldb : 1, transposeB() ? 1 : ldb,
                                   (int)N, (int)K, app.devicePtr(hB2_), (float*)app.devicePtr(hSB_),
                                   app.stream()) == ST_OK &&
               app.wait();
    }
    bool build(DeviceApp&) override { return validParams() && (dtype != DT_FP8 || packedCalc() == 1); }
    bool enqueue(DeviceApp& app) override {
        GemmArgs a;
        a.transA = transposeA();
        a.transB = transposeB();
        a.M = dimM();
        a.N = dimN();
        a.K = dimK();
        a.A = app.devicePtr(hA_);
        a.B = app.devicePtr(hB_);
        a.C = app.devicePtr(hC_);
        a.lda = transposeA() ? a.M : a.K;
        a.ldb = transposeB() ? a.K : a.N;
        a.ldc = a.N;
        a.strideA = a.M * a.K;
        a.strideB = a.K * a.N;
        a.strideC = a.M * a.N;
        a.batch = (int)packedCalc();
        a.alpha = generalizedMatmul() ? alpha_ : 1.0;
        a.beta = generalizedMatmul() ? beta_ : 0.0;
        if (dtype == DT_BF16) {
            a.A = app.devicePtr(hA2_);
            a.B = app.devicePtr(hB2_);
        } else if (dtype == DT_FP8) {
            a.A = app.devicePtr(hA2_);
            a.B = app.devicePtr(hB2_);
            a.transA = 0;
            a.transB = 1;
            a.lda = a.K;
            a.ldb = a.K;
            a.scaleA = (const float*)app.devicePtr(hSA_);
            a.scaleB = (const float*)app.devicePtr(hSB_);
        }
        const MfmaCfg c = cfg();
        return app.ctx().gemm(dtype, a, &c) == ST_OK;
    }
    bool syncOutput(DeviceApp& app) override { return app.syncFromDevice(hC_); }
    bool checkOutput(DeviceApp& app, bool) override { return f64_ ? check<double>(app) : check<float>(app); }
    std::ostream& print(std::ostream& os) const override {
        const MfmaCfg c = cfg();
        os << "// " << kernelName() << ": MFMA GEMM, " << c.waves_m << "x" << c.waves_n << " wavefronts, each "
           << c.tm << "x" << c.tn << " blocks of " << c.blk << "x" << c.blk << ", macro tile "
           << c.waves_m * c.tm * c.blk << " x " << c.waves_n * c.tn * c.blk << ", LDS depth " << c.bk
           << " (double-buffered)";
        if (c.wk > 1) os << ", each K tile split between " << c.wk << " wave groups";
        os << "\n//   M=" << dimM() << " N=" << dimN() << " K=" << dimK()
           << " transA=" << transposeA() << " transB=" << transposeB() << " GEMM=" << generalizedMatmul()
           << " packed=" << packedCalc() << " instantiated=" << (validParams() ? "yes" : "no") << "\n";
        return os;
    }
};

// ======================================================================
// matvec, reference geometry / native
// ======================================================================
template <typename T>
class KMatvec : public KernelBaseMatvec, public KernelTag {
    int vl_;
    int hA_ = -1, hX_ = -1, hY_ = -1;
    bool paranoid_ = false;
    T alpha_ = 1, beta_ = 1;
    std::vector<T> ref_;
    std::vector<double> ref64_, mag64_;
    double lastErr_ = 0;

public:
    KMatvec(int variant, int vl) : vl_(vl) {
        family = FAM_MATVEC;
        this->variant = variant;
        dtype = dtypeof<T>();
        if (variant == VAR_NATIVE) {
            xp_.addDigit(3);  // unroll {1,2,4}
        } else {
            xp_.addDigit(2);  // inlineMN
            if (variant == VAR_IMAGE) xp_.addDigit(2);  // globalID
        }
    }
    std::string kernelName() const override {
        std::ostringstream s;
        if (variant == VAR_NATIVE) s << "matvecnative" << nameof<T>();
        else s << (variant == VAR_IMAGE ? "matvecimage" : "matvecbuffer") << nameof<T>() << vl_;
        return s.str();
    }
    GemvCfg gcfg() const {
        static const int un[3] = {1, 2, 4};
        const auto d = extraParamDetail();
        return GemvCfg{(int)groupSize(), (int)blockHeight(), un[d.empty() ? 0 : std::min<size_t>(d[0], 2)]};
    }
    bool validParams() const override {
        if (variant != VAR_NATIVE) return KernelBaseMatvec::validParams();
        return packedCalc() > 0 && dimM() > 0 && dimN() > 0 && vectorLength() == 16 / sizeof(T) &&
               extraParam() < totalVariations() && gemv_cfg_ok(gcfg());
    }
    void paranoidCheck() override { paranoid_ = true; }
    double lastCheckError() const override { return lastErr_; }

    bool setArgs(DeviceApp& app, bool syncInput) override {
        const size_t P = packedCalc(), M = dimM(), N = dimN();
        if (hA_ < 0 || hX_ < 0 || hY_ < 0 || dimChanged() || packedChanged() || app.bytes(hY_) != P * M * sizeof(T)) {
            app.releaseBuffers();
            hA_ = app.createBuffer<T>(P * M * N, "matA", (T)1);
            hX_ = app.createBuffer<T>(P * N, "vecB", (T)1);
            hY_ = app.createBuffer<T>(P * M, "vecC", (T)0);
            if (hA_ < 0 || hX_ < 0 || hY_ < 0) return false;
        } else {
            if (syncInput && (!app.syncToDevice(hA_) || !app.syncToDevice(hX_))) return false;
            if (!app.clearBuffer(hY_)) return false;
        }
        alpha_ = (paranoid_ && generalizedMatvec()) ? checks::posrand<T>() : (T)1;
        beta_ = (paranoid_ && generalizedMatvec()) ? checks::posrand<T>() : (T)1;
        if (paranoid_) {
            T* A = app.bufferPtr<T>(hA_);
            T* X = app.bufferPtr<T>(hX_);
            T* Y = app.bufferPtr<T>(hY_);
            checks::fillrand(A, P * M * N);
            checks::fillrand(X, P * N);
            if (generalizedMatvec()) checks::fillrand(Y, P * M);
            if (!app.syncToDevice(hA_) || !app.syncToDevice(hX_) || !app.syncToDevice(hY_)) return false;
            if (variant != VAR_NATIVE) {
                ref_.assign(P * M, (T)0);
                checks::refMatvec<T>(P, generalizedMatvec(), transposeA(), M, N, alpha_, beta_, A, X,
                                     generalizedMatvec() ? Y : nullptr, ref_.data());
            } else {
                ref64_.assign(P * M, 0.0);
                mag64_.assign(P * M, 0.0);
                for (size_t p = 0; p < P; ++p)
                    for (size_t i = 0; i < M; ++i) {
                        double s = 0, m = 0;
                        for (size_t j = 0; j < N; ++j) {
                            const double a = transposeA() ? A[p * M * N + j * M + i] : A[p * M * N + i * N + j];
                            s += a * (double)X[p * N + j];
                            m += std::fabs(a) * std::fabs((double)X[p * N + j]);
                        }
                        const double y0 = generalizedMatvec() ? (double)Y[p * M + i] : 0.0;
                        const double al = generalizedMatvec() ? (double)alpha_ : 1.0;
                        const double be = generalizedMatvec() ? (double)beta_ : 0.0;
                        ref64_[p * M + i] = al * s + be * y0;
                        mag64_[p * M + i] = std::fabs(al) * m + std::fabs(be) * std::fabs(y0);
                    }
            }
        }
        return true;
    }
    bool build(DeviceApp&) override {
        if (!validParams()) return false;
        if (variant == VAR_NATIVE) return true;
        return groupSize() <= 1024 && vectorLength() == (size_t)vl_;
    }
    bool enqueue(DeviceApp& app) override {
        const size_t P = packedCalc(), M = dimM(), N = dimN();
        const bool g = generalizedMatvec();
        if (variant != VAR_NATIVE)
            return launch_gemv_compat(dtype, vl_, transposeA(), M, N, groupSize(), blockHeight(), (int)P, g,
                                      (double)alpha_, (double)beta_, app.devicePtr(hA_), app.devicePtr(hX_),
                                      app.devicePtr(hY_), app.stream()) == ST_OK;
        void* work = nullptr;
        const size_t wb = gemv_workspace_bytes(dtype, transposeA(), M, N);
        if (wb && app.ctx().workspace(wb, &work) != ST_OK) return false;
        const GemvCfg c = gcfg();
        for (size_t p = 0; p < P; ++p) {
            const T* A = (const T*)app.devicePtr(hA_) + p * M * N;
            const T* X = (const T*)app.devicePtr(hX_) + p * N;
            T* Y = (T*)app.devicePtr(hY_) + p * M;
            if (launch_gemv_native(dtype, transposeA(), M, N, g ? (double)alpha_ : 1.0, A, transposeA() ? M : N, X,
                                   g ? (double)beta_ : 0.0, Y, work, app.stream(), &c) != ST_OK)
                return false;
        }
        return true;
    }
    bool syncOutput(DeviceApp& app) override { return app.syncFromDevice(hY_); }
    bool checkOutput(DeviceApp& app, bool printOutput) override {
        const size_t n = packedCalc() * dimM();
        const T* Y = app.bufferPtr<T>(hY_);
        if (printOutput)
            for (size_t i = 0; i < std::min<size_t>(n, 64); ++i) std::cerr << Y[i] << (i + 1 < n ? " " : "\n");
        if (!paranoid_) return checks::checkArray(Y, n, (T)dimN());
        if (variant != VAR_NATIVE) {
            long double d = 0;
            const bool ok = checks::paranoidPass(Y, ref_.data(), n, &d);
            lastErr_ = (double)d;
            std::cerr << "absdiff: " << (double)d << "\t";
            return ok;
        }
        std::vector<double> got(Y, Y + n);
        lastErr_ = checks::componentwiseError(got.data(), ref64_.data(), mag64_.data(), n);
        std::cerr << "componentwise err: " << lastErr_ << "\t";
        return lastErr_ <= (sizeof(T) == 4 ? 1e-5 : 1e-12);
    }
    std::ostream& print(std::ostream& os) const override {
        os << "// " << kernelName() << ": ";
        if (variant == VAR_NATIVE) {
            const GemvCfg c = gcfg();
            os << (transposeA() ? "column-split partial sums + ordered reduction"
                                : "wavefront rows with butterfly reduction")
               << ", work-group " << c.wg << ", rows/wave " << c.rows << ", unroll " << c.unroll;
        } else {
            os << "gemv_compat_kernel, work-item owns " << blockHeight() << " rows, work-group " << groupSize();
        }
        os << "\n//   M=" << dimM() << " N=" << dimN() << " transA=" << transposeA()
           << " GEMV=" << generalizedMatvec() << " packed=" << packedCalc() << "\n";
        return os;
    }
};

// ======================================================================
// saxpy, reference geometry / native
// ======================================================================
template <typename T>
class KSaxpy : public KernelBaseSaxpy, public KernelTag {
    int vl_;
    int hX_ = -1, hY_ = -1, hZ_ = -1;
    bool paranoid_ = false;
    T alpha_ = 1;
    std::vector<T> ref_;
    double lastErr_ = 0;

public:
    KSaxpy(int variant, int vl) : vl_(vl) {
        family = FAM_SAXPY;
        this->variant = variant;
        dtype = dtypeof<T>();
        if (variant == VAR_NATIVE) {
            xp_.addDigit(2);  // non-temporal
        } else {
            xp_.addDigit(2);  // inlineMN
            if (variant == VAR_IMAGE) xp_.addDigit(2);  // globalID
        }
    }
    std::string kernelName() const override {
        std::ostringstream s;
        if (variant == VAR_NATIVE) s << "saxpynative" << nameof<T>();
        else s << (variant == VAR_IMAGE ? "saxpyimage" : "saxpybuffer") << nameof<T>() << vl_;
        return s.str();
    }
    AxpyCfg acfg() const {
        const auto d = extraParamDetail();
        return AxpyCfg{(int)groupHeight(), (int)blockHeight(), (int)(d.empty() ? 0 : d[0])};
    }
    bool validParams() const override {
        if (variant != VAR_NATIVE) return KernelBaseSaxpy::validParams();
        return packedCalc() > 0 && dimM() > 0 && dimN() == vectorLength() && vectorLength() == 16 / sizeof(T) &&
               groupWidth() == 0 && blockWidth() == 0 && extraParam() < totalVariations() && axpy_cfg_ok(acfg());
    }
    void paranoidCheck() override { paranoid_ = true; }
    double lastCheckError() const override { return lastErr_; }
    size_t bufferSize() const { return packedCalc() * dimM() * dimN(); }

    bool setArgs(DeviceApp& app, bool syncInput) override {
        const size_t n = bufferSize();
        if (hX_ < 0 || hY_ < 0 || hZ_ < 0 || dimChanged() || packedChanged() || app.bytes(hZ_) != n * sizeof(T)) {
            app.releaseBuffers();
            hX_ = app.createBuffer<T>(n, "X", (T)1);
            hY_ = app.createBuffer<T>(n, "Y", (T)1);
            hZ_ = app.createBuffer<T>(n, "Z", (T)0);
            if (hX_ < 0 || hY_ < 0 || hZ_ < 0) return false;
        } else {
            if (syncInput && (!app.syncToDevice(hX_) || !app.syncToDevice(hY_))) return false;
            if (!app.clearBuffer(hZ_)) return false;
        }
        alpha_ = paranoid_ ? checks::posrand<T>() : (T)1;
        if (paranoid_) {
            T* X = app.bufferPtr<T>(hX_);
            T* Y = app.bufferPtr<T>(hY_);
            checks::fillrand(X, n);
            checks::fillrand(Y, n);
            if (!app.syncToDevice(hX_) || !app.syncToDevice(hY_)) return false;
            ref_.assign(n, (T)0);
            checks::refSaxpy<T>(n, alpha_, X, Y, ref_.data());
        }
        return true;
    }
    bool build(DeviceApp&) override {
        if (!validParams()) return false;
        if (variant == VAR_NATIVE) return true;
        const size_t threads = groupHeight() * (groupWidth() ? groupWidth() : 1);
        return threads >= 1 && threads <= 1024 && vectorLength() == (size_t)vl_;
    }
    bool enqueue(DeviceApp& app) override {
        if (variant == VAR_NATIVE) {
            const AxpyCfg c = acfg();
            return launch_saxpy_native(dtype, bufferSize(), (double)alpha_, app.devicePtr(hX_), app.devicePtr(hY_),
                                       app.devicePtr(hZ_), app.stream(), &c) == ST_OK;
        }
        return launch_saxpy_compat(dtype, vl_, dimM(), dimN(), groupHeight(), groupWidth(), blockHeight(),
                                   blockWidth(), (int)packedCalc(), (double)alpha_, app.devicePtr(hX_),
                                   app.devicePtr(hY_), app.devicePtr(hZ_), app.stream()) == ST_OK;
    }
    bool syncOutput(DeviceApp& app) override { return app.syncFromDevice(hZ_); }
    bool checkOutput(DeviceApp& app, bool printOutput) override {
        const size_t n = bufferSize();
        const T* Z = app.bufferPtr<T>(hZ_);
        if (printOutput)
            for (size_t i = 0; i < std::min<size_t>(n, 64); ++i) std::cerr << Z[i] << (i + 1 < n ? " " : "\n");
        if (!paranoid_) return checks::checkArray(Z, n, (T)(1 * 1 + 1));
        long double d = 0;
        const bool ok = checks::paranoidPass(Z, ref_.data(), n, &d);
        lastErr_ = (double)d;
        std::cerr << "absdiff: " << (double)d << "\t";
        return ok;
    }
    std::ostream& print(std::ostream& os) const override {
        os << "// " << kernelName() << ": ";
        if (variant == VAR_NATIVE) {
            const AxpyCfg c = acfg();
            os << "streaming z = alpha*x + y, work-group " << c.wg << ", 16-byte vectors/work-item " << c.unroll
               << (c.nontemporal ? ", non-temporal" : ", cached");
        } else {
            os << "saxpy_compat_kernel, " << (dimN() == (size_t)vl_ ? "1-D" : "2-D") << " work-group "
               << groupHeight() << (groupWidth() ? " x " + std::to_string(groupWidth()) : std::string())
               << ", blocking " << blockHeight() << " x " << blockWidth();
        }
        os << "\n//   M=" << dimM() << " N=" << dimN() << " packed=" << packedCalc() << "\n";
        return os;
    }
};

// ======================================================================
KernelInterface* make_kernel(int family, int dtype, int vl, int variant) {
    const bool f = dtype == DT_F32, d = dtype == DT_F64;
    switch (family) {
        case FAM_MATMUL:
            if (variant == VAR_BUFFER && (vl == 1 || vl == 2 || vl == 4)) {
                if (f) return new KMatmulCompat<float>(VAR_BUFFER, vl);
                if (d) return new KMatmulCompat<double>(VAR_BUFFER, vl);
            }
            if (variant == VAR_IMAGE) {
                if (f && vl == 4) return new KMatmulCompat<float>(VAR_IMAGE, 4);
                if (d && vl == 2) return new KMatmulCompat<double>(VAR_IMAGE, 2);
            }
            if (variant == VAR_MFMA && dtype >= DT_F32 && dtype <= DT_FP8) return new KMatmulMfma(dtype);
            break;
        case FAM_MATVEC:
            if (variant == VAR_BUFFER && (vl == 1 || vl == 2 || vl == 4)) {
                if (f) return new KMatvec<float>(VAR_BUFFER, vl);
                if (d) return new KMatvec<double>(VAR_BUFFER, vl);
            }
            if (variant == VAR_IMAGE) {
                if (f && vl == 4) return new KMatvec<float>(VAR_IMAGE, 4);
                if (d && vl == 2) return new KMatvec<double>(VAR_IMAGE, 2);
            }
            if (variant == VAR_NATIVE) {
                if (f) return new KMatvec<float>(VAR_NATIVE, 4);
                if (d) return new KMatvec<double>(VAR_NATIVE, 2);
            }
            break;
        case FAM_SAXPY:
            if (variant == VAR_BUFFER && (vl == 1 || vl == 2 || vl == 4)) {
                if (f) return new KSaxpy<float>(VAR_BUFFER, vl);
                if (d) return new KSaxpy<double>(VAR_BUFFER, vl);
            }
            if (variant == VAR_IMAGE) {
                if (f && vl == 4) return new KSaxpy<float>(VAR_IMAGE, 4);
                if (d && vl == 2) return new KSaxpy<double>(VAR_IMAGE, 2);
            }
            if (variant == VAR_NATIVE) {
                if (f) return new KSaxpy<float>(VAR_NATIVE, 4);
                if (d) return new KSaxpy<double>(VAR_NATIVE, 2);
            }
            break;
    }
    return nullptr;
}

static KernelTag* tag(KernelInterface* k) { return dynamic_cast<KernelTag*>(k); }
KernelBaseMatmul* as_matmul(KernelInterface* k) { return dynamic_cast<KernelBaseMatmul*>(k); }
KernelBaseMatvec* as_matvec(KernelInterface* k) { return dynamic_cast<KernelBaseMatvec*>(k); }
KernelBaseSaxpy* as_saxpy(KernelInterface* k) { return dynamic_cast<KernelBaseSaxpy*>(k); }
int kernel_family(KernelInterface* k) { return tag(k) ? tag(k)->family : -1; }
int kernel_variant(KernelInterface* k) { return tag(k) ? tag(k)->variant : -1; }
int kernel_dtype(KernelInterface* k) { return tag(k) ? tag(k)->dtype : -1; }

}  // namespace gat
