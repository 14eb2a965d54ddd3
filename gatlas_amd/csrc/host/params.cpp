#include "params.hpp"

namespace gat {

std::ostream& operator<<(std::ostream& os, const KernelInterface& k) { return k.print(os); }

size_t ExtraParam::addDigit(size_t radix) {
    radix_.push_back(radix);
    digits_.push_back(0);
    total_ *= radix;
    return radix_.size() - 1;
}

void ExtraParam::setExtraParameter(size_t value) {
    value_ = value;
    size_t place = 1;
    for (size_t i = 0; i < radix_.size(); ++i) {
        digits_[i] = (value / place) % radix_[i];
        place *= radix_[i];
    }
}

// ---------------------------------------------------------------- matmul
KernelBaseMatmul::KernelBaseMatmul() {}

void KernelBaseMatmul::setPackedCalc(size_t n) {
    packedChanged_ = (n != packed_);
    packed_ = n;
}
void KernelBaseMatmul::setGeneralizedMatmul(bool gemm) {
    gemmChanged_ = (gemm != gemm_);
    gemm_ = gemm;
}
void KernelBaseMatmul::setMatrixDimensions(size_t M, size_t N, size_t K) {
    dimChanged_ = (M != M_) || (N != N_) || (K != K_);
    M_ = M;
    N_ = N;
    K_ = K;
}
void KernelBaseMatmul::setDataLayout(bool a, bool b) {
    layoutChanged_ = (a != tA_) || (b != tB_);
    tA_ = a;
    tB_ = b;
}
void KernelBaseMatmul::setWorkGroup(size_t h, size_t w) {
    gH_ = h;
    gW_ = w;
}
void KernelBaseMatmul::setInnerBlocking(size_t h, size_t w) {
    bh_ = h;
    bw_ = w;
}

bool KernelBaseMatmul::validParams() const {
    const size_t VL = bw_;
    if (packed_ == 0 || VL == 0 || gH_ == 0 || gW_ == 0 || bh_ == 0) return false;
    return M_ % VL == 0 && N_ % VL == 0 && K_ % VL == 0 &&       //
           M_ % (gH_ * bh_) == 0 && N_ % (gW_ * bw_) == 0 &&     //
           K_ % (gH_ * VL) == 0 &&                               //
           gH_ * bh_ <= M_ && gW_ * bw_ <= N_ && gH_ * VL <= K_ &&  //
           (tA_ ? (bh_ % VL == 0) : true) &&                     //
           extraParam() < totalVariations();
}

bool KernelBaseMatmul::getParams(std::vector<size_t>& p) const {
    if (!validParams()) return false;
    p = {packed_, (size_t)gemm_, M_, N_, K_, (size_t)tA_, (size_t)tB_, gH_, gW_, bh_, bw_, extraParam()};
    return true;
}

void KernelBaseMatmul::setParams(const std::vector<size_t>& p) {
    size_t i = 0;
    setPackedCalc(p[i++]);
    setGeneralizedMatmul(p[i++] != 0);
    const size_t M = p[i++], N = p[i++], K = p[i++];
    setMatrixDimensions(M, N, K);
    const bool a = p[i++] != 0, b = p[i++] != 0;
    setDataLayout(a, b);
    const size_t h = p[i++], w = p[i++];
    setWorkGroup(h, w);
    const size_t bh = p[i++], bw = p[i++];
    setInnerBlocking(bh, bw);
    setExtraParameter(p[i++]);
}

std::vector<size_t> KernelBaseMatmul::globalWorkItems() const {
    return {bw_ ? N_ / bw_ : 0, bh_ ? M_ / bh_ : 0};
}
std::vector<size_t> KernelBaseMatmul::localWorkItems() const { return {gW_, gH_}; }

size_t KernelBaseMatmul::numberFlops() const {
    if (gemm_) return packed_ * 2 * M_ * N_ * (K_ + 1);
    return packed_ * M_ * N_ * (2 * K_ - 1);
}

// ---------------------------------------------------------------- matvec
void KernelBaseMatvec::setPackedCalc(size_t n) {
    packedChanged_ = (n != packed_);
    packed_ = n;
}
void KernelBaseMatvec::setGeneralizedMatvec(bool g) {
    gemvChanged_ = (g != gemv_);
    gemv_ = g;
}
void KernelBaseMatvec::setMatrixDimensions(size_t M, size_t N) {
    dimChanged_ = (M != M_) || (N != N_);
    M_ = M;
    N_ = N;
}
void KernelBaseMatvec::setDataLayout(bool a) {
    layoutChanged_ = (a != tA_);
    tA_ = a;
}
void KernelBaseMatvec::setInnerBlocking(size_t h, size_t vl) {
    bh_ = h;
    vl_ = vl;
}

bool KernelBaseMatvec::validParams() const {
    if (packed_ == 0 || vl_ == 0 || gs_ == 0 || bh_ == 0) return false;
    return M_ % vl_ == 0 && N_ % vl_ == 0 &&                       //
           M_ % (gs_ * bh_) == 0 && N_ % (gs_ * vl_) == 0 &&       //
           gs_ * bh_ <= M_ && gs_ * vl_ <= N_ &&                   //
           bh_ % vl_ == 0 && extraParam() < totalVariations();
}

bool KernelBaseMatvec::getParams(std::vector<size_t>& p) const {
    if (!validParams()) return false;
    p = {packed_, (size_t)gemv_, M_, N_, (size_t)tA_, gs_, bh_, vl_, extraParam()};
    return true;
}

void KernelBaseMatvec::setParams(const std::vector<size_t>& p) {
    size_t i = 0;
    setPackedCalc(p[i++]);
    setGeneralizedMatvec(p[i++] != 0);
    const size_t M = p[i++], N = p[i++];
    setMatrixDimensions(M, N);
    setDataLayout(p[i++] != 0);
    setWorkGroup(p[i++]);
    const size_t bh = p[i++], vl = p[i++];
    setInnerBlocking(bh, vl);
    setExtraParameter(p[i++]);
}

std::vector<size_t> KernelBaseMatvec::globalWorkItems() const { return {bh_ ? M_ / bh_ : 0}; }
std::vector<size_t> KernelBaseMatvec::localWorkItems() const { return {gs_}; }

size_t KernelBaseMatvec::numberFlops() const {
    if (gemv_) return packed_ * 2 * M_ * (N_ + 1);
    return packed_ * M_ * (2 * N_ - 1);
}

// ----------------------------------------------------------------- saxpy
void KernelBaseSaxpy::setPackedCalc(size_t n) {
    packedChanged_ = (n != packed_);
    packed_ = n;
}
void KernelBaseSaxpy::setSaxpyDimensions(size_t M, size_t N) {
    dimChanged_ = (M != M_) || (N != N_);
    M_ = M;
    N_ = N;
}

bool KernelBaseSaxpy::validParams() const {
    if (packed_ == 0 || vl_ == 0 || gH_ == 0 || bh_ == 0) return false;
    return M_ % vl_ == 0 && N_ % vl_ == 0 &&                                    //
           M_ % (gH_ * bh_) == 0 &&                                             //
           ((gW_ && bw_) ? N_ % (gW_ * bw_) == 0 : true) &&                     //
           gH_ * bh_ <= M_ && gW_ * bw_ <= N_ &&                                //
           bw_ % vl_ == 0 && extraParam() < totalVariations();
}

bool KernelBaseSaxpy::getParams(std::vector<size_t>& p) const {
    if (!validParams()) return false;
    p = {packed_, M_, N_, gH_, gW_, bh_, bw_, vl_, extraParam()};
    return true;
}

void KernelBaseSaxpy::setParams(const std::vector<size_t>& p) {
    size_t i = 0;
    setPackedCalc(p[i++]);
    const size_t M = p[i++], N = p[i++];
    setSaxpyDimensions(M, N);
    const size_t h = p[i++], w = p[i++];
    setWorkGroup(h, w);
    const size_t bh = p[i++], bw = p[i++];
    setInnerBlocking(bh, bw);
    setVectorLength(p[i++]);
    setExtraParameter(p[i++]);
}

std::vector<size_t> KernelBaseSaxpy::globalWorkItems() const {
    std::vector<size_t> d{bh_ ? M_ / bh_ : 0};
    if (bw_) d.push_back(N_ / bw_);
    return d;
}
std::vector<size_t> KernelBaseSaxpy::localWorkItems() const {
    std::vector<size_t> d{gH_};
    if (gW_) d.push_back(gW_);
    return d;
}

size_t KernelBaseSaxpy::numberFlops() const { return packed_ * 2 * M_ * N_; }

}  // namespace gat
