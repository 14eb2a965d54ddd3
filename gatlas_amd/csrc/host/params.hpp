// Parameter interface of the three kernel families, with the semantics of
// the reference's KernelBaseMatmul / KernelBaseMatvec / KernelBaseSaxpy
// (validity predicate, 12/9/9-element parameter vectors, work-item shapes,
// flop counts, mixed-radix extra parameter).  Pure host code.
#pragma once

#include <cstddef>
#include <ostream>
#include <string>
#include <vector>

namespace gat {

class DeviceApp;

// The plugin boundary (GatlasBenchmark.hpp:28-58 KernelInterface), with the
// OpenCL program build replaced by resolving a compiled HIP instantiation
// (build) and the NDRange enqueue replaced by a HIP launch (enqueue).
struct KernelInterface {
    virtual ~KernelInterface() {}
    virtual std::string kernelName() const = 0;
    virtual size_t numberFlops() const = 0;
    virtual void setParams(const std::vector<size_t>& params) = 0;
    virtual std::vector<size_t> extraParamDetail() const = 0;
    virtual bool setArgs(DeviceApp& app, bool syncInput) = 0;
    virtual bool syncOutput(DeviceApp& app) = 0;
    virtual bool checkOutput(DeviceApp& app, bool printOutput = false) = 0;
    virtual void paranoidCheck() = 0;
    virtual std::vector<size_t> globalWorkItems() const = 0;
    virtual std::vector<size_t> localWorkItems() const = 0;
    virtual std::ostream& print(std::ostream& os) const = 0;
    // true when the current parameters name a compiled instantiation
    virtual bool build(DeviceApp& app) = 0;
    // launch once on the app's stream (asynchronous)
    virtual bool enqueue(DeviceApp& app) = 0;
    // last check's absdiff (paranoid rule) or componentwise error (FP64 rule)
    virtual double lastCheckError() const { return 0.0; }
};

std::ostream& operator<<(std::ostream& os, const KernelInterface& k);

// Mixed-radix extra parameter (KernelBaseMatmul.cpp:146-202): digit i of
// value x is (x / prod_{j<i} radix_j) % radix_i; the first registered digit
// is the least significant; totalVariations is the product of the radices.
class ExtraParam {
public:
    size_t addDigit(size_t radix);
    void setExtraParameter(size_t value);
    size_t extraParam() const { return value_; }
    size_t totalVariations() const { return total_; }
    size_t digit(size_t i) const { return digits_[i]; }
    std::vector<size_t> extraParamDetail() const { return digits_; }

private:
    size_t value_ = 0, total_ = 1;
    std::vector<size_t> radix_, digits_;
};

// ---------------------------------------------------------------- matmul
class KernelBaseMatmul : public KernelInterface {
public:
    KernelBaseMatmul();

    void setUseAttrAutoVec(bool v) { autoVec_ = v; }
    bool getUseAttrAutoVec() const { return autoVec_; }
    void setPackedCalc(size_t n);
    void setGeneralizedMatmul(bool gemm);
    void setMatrixDimensions(size_t M, size_t N, size_t K);
    void setDataLayout(bool transA, bool transB);
    void setWorkGroup(size_t height, size_t width);
    void setWorkGroup(size_t sz) { setWorkGroup(sz, sz); }
    void setInnerBlocking(size_t height, size_t width);
    void setExtraParameter(size_t v) { xp_.setExtraParameter(v); }

    size_t packedCalc() const { return packed_; }
    bool generalizedMatmul() const { return gemm_; }
    size_t dimM() const { return M_; }
    size_t dimN() const { return N_; }
    size_t dimK() const { return K_; }
    bool transposeA() const { return tA_; }
    bool transposeB() const { return tB_; }
    size_t groupHeight() const { return gH_; }
    size_t groupWidth() const { return gW_; }
    size_t groupSize() const { return gH_ == gW_ ? gW_ : 0; }
    size_t blockHeight() const { return bh_; }
    size_t blockWidth() const { return bw_; }
    size_t extraParam() const { return xp_.extraParam(); }
    size_t totalVariations() const { return xp_.totalVariations(); }
    std::vector<size_t> extraParamDetail() const override { return xp_.extraParamDetail(); }

    bool dimChanged() const { return dimChanged_; }
    bool layoutChanged() const { return layoutChanged_; }
    bool gemmChanged() const { return gemmChanged_; }
    bool packedChanged() const { return packedChanged_; }

    // KernelBaseMatmul.cpp:263-291 (overridable: MFMA kernels validate
    // against their instantiated tile table instead of the blocking rule)
    virtual bool validParams() const;
    bool getParams(std::vector<size_t>& params) const;
    void setParams(const std::vector<size_t>& params) override;
    std::vector<size_t> globalWorkItems() const override;
    std::vector<size_t> localWorkItems() const override;
    size_t numberFlops() const override;

protected:
    ExtraParam xp_;

private:
    bool autoVec_ = true;
    size_t packed_ = 1;
    bool gemm_ = false;
    size_t M_ = 0, N_ = 0, K_ = 0;
    bool tA_ = false, tB_ = false;
    size_t gH_ = 0, gW_ = 0, bh_ = 0, bw_ = 0;
    bool dimChanged_ = false, layoutChanged_ = false, gemmChanged_ = false, packedChanged_ = false;
};

// ---------------------------------------------------------------- matvec
class KernelBaseMatvec : public KernelInterface {
public:
    void setUseAttrAutoVec(bool v) { autoVec_ = v; }
    void setPackedCalc(size_t n);
    void setGeneralizedMatvec(bool gemv);
    void setMatrixDimensions(size_t M, size_t N);
    void setDataLayout(bool transA);
    void setWorkGroup(size_t sz) { gs_ = sz; }
    void setInnerBlocking(size_t height, size_t vectorLength);
    void setExtraParameter(size_t v) { xp_.setExtraParameter(v); }

    size_t packedCalc() const { return packed_; }
    bool generalizedMatvec() const { return gemv_; }
    size_t dimM() const { return M_; }
    size_t dimN() const { return N_; }
    bool transposeA() const { return tA_; }
    size_t groupSize() const { return gs_; }
    size_t blockHeight() const { return bh_; }
    size_t vectorLength() const { return vl_; }
    size_t extraParam() const { return xp_.extraParam(); }
    size_t totalVariations() const { return xp_.totalVariations(); }
    std::vector<size_t> extraParamDetail() const override { return xp_.extraParamDetail(); }
    bool dimChanged() const { return dimChanged_; }
    bool packedChanged() const { return packedChanged_; }

    // KernelBaseMatvec.cpp:236-257
    virtual bool validParams() const;
    bool getParams(std::vector<size_t>& params) const;
    void setParams(const std::vector<size_t>& params) override;
    std::vector<size_t> globalWorkItems() const override;
    std::vector<size_t> localWorkItems() const override;
    size_t numberFlops() const override;

protected:
    ExtraParam xp_;

private:
    bool autoVec_ = true;
    size_t packed_ = 1;
    bool gemv_ = false;
    size_t M_ = 0, N_ = 0;
    bool tA_ = false;
    size_t gs_ = 0, bh_ = 0, vl_ = 0;
    bool dimChanged_ = false, layoutChanged_ = false, gemvChanged_ = false, packedChanged_ = false;
};

// ----------------------------------------------------------------- saxpy
class KernelBaseSaxpy : public KernelInterface {
public:
    void setUseAttrAutoVec(bool v) { autoVec_ = v; }
    void setPackedCalc(size_t n);
    void setSaxpyDimensions(size_t M, size_t N);
    void setWorkGroup(size_t height, size_t width) {
        gH_ = height;
        gW_ = width;
    }
    void setInnerBlocking(size_t height, size_t width) {
        bh_ = height;
        bw_ = width;
    }
    void setVectorLength(size_t vl) { vl_ = vl; }
    void setExtraParameter(size_t v) { xp_.setExtraParameter(v); }

    size_t packedCalc() const { return packed_; }
    size_t dimM() const { return M_; }
    size_t dimN() const { return N_; }
    size_t groupHeight() const { return gH_; }
    size_t groupWidth() const { return gW_; }
    size_t blockHeight() const { return bh_; }
    size_t blockWidth() const { return bw_; }
    size_t vectorLength() const { return vl_; }
    size_t extraParam() const { return xp_.extraParam(); }
    size_t totalVariations() const { return xp_.totalVariations(); }
    std::vector<size_t> extraParamDetail() const override { return xp_.extraParamDetail(); }
    bool dimChanged() const { return dimChanged_; }
    bool packedChanged() const { return packedChanged_; }

    // KernelBaseSaxpy.cpp:210-231
    virtual bool validParams() const;
    bool getParams(std::vector<size_t>& params) const;
    void setParams(const std::vector<size_t>& params) override;
    std::vector<size_t> globalWorkItems() const override;
    std::vector<size_t> localWorkItems() const override;
    size_t numberFlops() const override;

protected:
    ExtraParam xp_;

private:
    bool autoVec_ = true;
    size_t packed_ = 1;
    size_t M_ = 0, N_ = 0;
    size_t gH_ = 0, gW_ = 0, bh_ = 0, bw_ = 0, vl_ = 0;
    bool dimChanged_ = false, packedChanged_ = false;
};

}  // namespace gat
