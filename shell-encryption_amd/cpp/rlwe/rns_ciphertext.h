// RnsRlweCiphertext (rns/rns_ciphertext.h:44-631) and RnsBgvCiphertext
// (rns/rns_bgv_ciphertext.h:35-175): the ciphertext operations on the hot
// path (tensor product, modulus switch, substitution, the absorb updates of
// the "b" part, eager and lazy), each a batched libshellgpu launch over the
// components.
#ifndef SHG_RLWE_RNS_CIPHERTEXT_H_
#define SHG_RLWE_RNS_CIPHERTEXT_H_

#include <string>
#include <vector>

#include "rlwe/lazy_rns_polynomial.h"
#include "rlwe/rns_gadget.h"
#include "rlwe/rns_polynomial.h"

namespace rlwe {

template <typename ModularInt>
class RnsRlweCiphertext {
 public:
  using ModuliVec = std::vector<const PrimeModulus<ModularInt>*>;

  explicit RnsRlweCiphertext(std::vector<RnsPolynomial<ModularInt>> components, ModuliVec moduli,
                             int power_of_s, double error,
                             const RnsErrorParams<ModularInt>* error_params)
      : components_(std::move(components)), moduli_(std::move(moduli)),
        error_params_(error_params), power_of_s_(power_of_s), error_(error) {}

  Status NegateInPlace() {
    for (auto& c : components_) RLWE_RETURN_IF_ERROR(c.NegateInPlace(moduli_));
    return OkStatus();
  }
  Status AddInPlace(const RnsRlweCiphertext& that) {
    if (Degree() != that.Degree()) return InvalidArgumentError("`that` has a mismatched degree.");
    if (Level() != that.Level()) return InvalidArgumentError("`that` has a mismatched level.");
    if (PowerOfS() != that.PowerOfS())
      return InvalidArgumentError("`that` is encrypted with a different key power.");
    for (size_t i = 0; i < components_.size(); ++i)
      RLWE_RETURN_IF_ERROR(components_[i].AddInPlace(that.components_[i], moduli_));
    error_ += that.error_;
    return OkStatus();
  }

  // rns/rns_ciphertext.h:501-524: one tensor-product launch.
  StatusOr<RnsRlweCiphertext> Mul(const RnsRlweCiphertext& that) const {
    if (Level() != that.Level()) return InvalidArgumentError("`that` has a mismatched level.");
    if (PowerOfS() != that.PowerOfS())
      return InvalidArgumentError("Ciphertexts must be encrypted with the same key power.");
    const int la = Len(), lb = that.Len(), L = NumModuli();
    for (const auto& c : components_)
      if (!c.IsNttForm()) return InvalidArgumentError("RNS polynomial `this` must be in NTT form.");
    for (const auto& c : that.components_)
      if (!c.IsNttForm()) return InvalidArgumentError("RNS polynomial `a` must be in NTT form.");
    const size_t words = (size_t)L << LogN();
    device::Buffer a(words * la), b(words * lb), out(words * (la + lb - 1));
    for (int i = 0; i < la; ++i)
      RLWE_RETURN_IF_ERROR(device::CopyWords(a.data() + i * words, components_[i].DeviceData(), words));
    for (int j = 0; j < lb; ++j)
      RLWE_RETURN_IF_ERROR(device::CopyWords(b.data() + j * words, that.components_[j].DeviceData(), words));
    auto ids = DevicePrimes<ModularInt>(ModuliSpan<ModularInt>(moduli_.data(), moduli_.size()));
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_tensor_product(
        device::Context(), a.data(), la, b.data(), lb, out.data(), 1, L, ids.data())));
    std::vector<RnsPolynomial<ModularInt>> result;
    for (int m = 0; m < la + lb - 1; ++m)
      result.push_back(RnsPolynomial<ModularInt>::FromDevice(LogN(), L, true, out.data() + m * words));
    return RnsRlweCiphertext(std::move(result), moduli_, power_of_s_, error_ * that.error_,
                             error_params_);
  }

  // rns/rns_ciphertext.h:309-330: this[0] += ctxt[0] * ptxt, the "a" part
  // left alone (set later with SetPadComponent).
  Status FusedAbsorbAddInPlaceWithoutPad(const RnsRlweCiphertext& ctxt,
                                         const RnsPolynomial<ModularInt>& ptxt) {
    if (Degree() != 1 || Degree() != ctxt.Degree())
      return InvalidArgumentError(
          "`FusedAbsorbAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
    RLWE_RETURN_IF_ERROR(CheckAbsorbOperand(ctxt));
    RLWE_RETURN_IF_ERROR(components_[0].FusedMulAddInPlace(ctxt.components_[0], ptxt, moduli_));
    error_ += ctxt.error_ * error_params_->B_plaintext();
    return OkStatus();
  }

  // rns/rns_ciphertext.h:338-368: the same into a lazy "b" part, which
  // MergeLazyOperations folds back into the components.
  Status FusedAbsorbAddInPlaceWithoutPadLazily(const RnsRlweCiphertext& ctxt,
                                               const RnsPolynomial<ModularInt>& ptxt) {
    if (ctxt.Degree() != 1)
      return InvalidArgumentError(
          "`FusedAbsorbAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
    RLWE_RETURN_IF_ERROR(CheckAbsorbOperand(ctxt));
    if (lazy_components_.empty()) {
      RLWE_ASSIGN_OR_RETURN(auto lazy, LazyRnsPolynomial<ModularInt>::Create(
                                           ctxt.components_[0], ptxt, Moduli()));
      lazy_components_.push_back(std::move(lazy));
    } else {
      RLWE_RETURN_IF_ERROR(
          lazy_components_[0].FusedMulAddInPlace(ctxt.components_[0], ptxt, Moduli()));
    }
    error_ += ctxt.error_ * error_params_->B_plaintext();
    return OkStatus();
  }

  // rns/rns_ciphertext.h:370-406: (ctxt1[0] + ctxt2[0]) * ptxt, lazily.
  Status FusedAbsorbSumAddInPlaceWithoutPadLazily(const RnsRlweCiphertext& ctxt1,
                                                  const RnsRlweCiphertext& ctxt2,
                                                  const RnsPolynomial<ModularInt>& ptxt) {
    if (ctxt1.Degree() != 1 || ctxt2.Degree() != 1)
      return InvalidArgumentError(
          "`FusedAbsorbSumAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
    RLWE_RETURN_IF_ERROR(CheckAbsorbOperands(ctxt1, ctxt2));
    if (lazy_components_.empty()) {
      RLWE_ASSIGN_OR_RETURN(auto sum, ctxt1.components_[0].Add(ctxt2.components_[0], Moduli()));
      RLWE_ASSIGN_OR_RETURN(auto lazy, LazyRnsPolynomial<ModularInt>::Create(sum, ptxt, Moduli()));
      lazy_components_.push_back(std::move(lazy));
    } else {
      RLWE_RETURN_IF_ERROR(lazy_components_[0].FusedMulSumAddInPlace(
          ctxt1.components_[0], ctxt2.components_[0], ptxt, Moduli()));
    }
    error_ += (ctxt1.error_ + ctxt2.error_) * error_params_->B_plaintext();
    return OkStatus();
  }

  // rns/rns_ciphertext.h:408-443: (ctxt1[0] - ctxt2[0]) * ptxt, lazily.  As
  // in the reference, the call that creates the lazy part leaves the error
  // unchanged.
  Status FusedAbsorbDiffAddInPlaceWithoutPadLazily(const RnsRlweCiphertext& ctxt1,
                                                   const RnsRlweCiphertext& ctxt2,
                                                   const RnsPolynomial<ModularInt>& ptxt) {
    if (ctxt1.Degree() != 1 || ctxt2.Degree() != 1)
      return InvalidArgumentError(
          "`FusedAbsorbDiffAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
    RLWE_RETURN_IF_ERROR(CheckAbsorbOperands(ctxt1, ctxt2));
    if (lazy_components_.empty()) {
      RLWE_ASSIGN_OR_RETURN(auto diff, ctxt1.components_[0].Sub(ctxt2.components_[0], Moduli()));
      RLWE_ASSIGN_OR_RETURN(auto lazy,
                            LazyRnsPolynomial<ModularInt>::Create(diff, ptxt, Moduli()));
      lazy_components_.push_back(std::move(lazy));
      return OkStatus();
    }
    RLWE_RETURN_IF_ERROR(lazy_components_[0].FusedMulDifferenceAddInPlace(
        ctxt1.components_[0], ctxt2.components_[0], ptxt, Moduli()));
    error_ += (ctxt1.error_ + ctxt2.error_) * error_params_->B_plaintext();
    return OkStatus();
  }

  // rns/rns_ciphertext.h:446-479.
  Status MergeLazyOperations() {
    if (lazy_components_.empty())
      return FailedPreconditionError("The ciphertext was not in lazy mode.");
    for (size_t i = 0; i < lazy_components_.size() && i < components_.size(); ++i) {
      RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> ci, lazy_components_[i].Export(Moduli()));
      RLWE_RETURN_IF_ERROR(components_[i].AddInPlace(ci, moduli_));
    }
    for (size_t i = components_.size(); i < lazy_components_.size(); ++i) {
      RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> ci, lazy_components_[i].Export(Moduli()));
      components_.push_back(std::move(ci));
    }
    lazy_components_.clear();
    return OkStatus();
  }

  // rns/rns_ciphertext.h:483-495: the "a" part; a ciphertext of degree 0
  // (lazy absorbs into an empty ciphertext) gains it as a new component.
  Status SetPadComponent(RnsPolynomial<ModularInt> new_a) {
    if (Degree() > 1)
      return InvalidArgumentError("`SetPadComponent` only applies to degree 1 ciphertext.");
    if (Degree() == 1)
      components_[1] = std::move(new_a);
    else
      components_.push_back(std::move(new_a));
    return OkStatus();
  }

  // rns/rns_ciphertext.h:70-98: each component through
  // RnsPolynomial::Serialize / Deserialize (bit-packed on the device).
  static StatusOr<RnsRlweCiphertext> Deserialize(const SerializedRnsRlweCiphertext& serialized,
                                                 ModuliVec moduli,
                                                 const RnsErrorParams<ModularInt>* error_params) {
    if (error_params == nullptr) return InvalidArgumentError("`error_params` must not be null.");
    std::vector<RnsPolynomial<ModularInt>> components;
    for (int i = 0; i < serialized.components_size(); ++i) {
      RLWE_ASSIGN_OR_RETURN(auto c, RnsPolynomial<ModularInt>::Deserialize(
                                        serialized.components(i),
                                        ModuliSpan<ModularInt>(moduli.data(), moduli.size())));
      components.push_back(std::move(c));
    }
    return RnsRlweCiphertext(std::move(components), std::move(moduli), serialized.power_of_s(),
                             serialized.error(), error_params);
  }
  StatusOr<SerializedRnsRlweCiphertext> Serialize() const {
    SerializedRnsRlweCiphertext serialized;
    for (const auto& c : components_) {
      RLWE_ASSIGN_OR_RETURN(*serialized.add_components(), c.Serialize(Moduli()));
    }
    serialized.set_power_of_s(power_of_s_);
    serialized.set_error(error_);
    return serialized;
  }

  Status ConvertToNttForm() {
    for (auto& c : components_) RLWE_RETURN_IF_ERROR(c.ConvertToNttForm(moduli_));
    return OkStatus();
  }
  Status ConvertToCoeffForm() {
    for (auto& c : components_) RLWE_RETURN_IF_ERROR(c.ConvertToCoeffForm(moduli_));
    return OkStatus();
  }

  int Len() const { return (int)components_.size(); }
  int Degree() const { return (int)components_.size() - 1; }
  int Level() const { return (int)moduli_.size() - 1; }
  StatusOr<RnsPolynomial<ModularInt>> Component(int index) const {
    if (index < 0 || index >= (int)components_.size())
      return InvalidArgumentError("Index out of range.");
    return components_[index];
  }
  int LogN() const { return components_[0].LogN(); }
  int NumCoeffs() const { return components_[0].NumCoeffs(); }
  int NumModuli() const { return (int)moduli_.size(); }
  ModuliSpan<ModularInt> Moduli() const { return ModuliSpan<ModularInt>(moduli_.data(), moduli_.size()); }
  const ModuliVec& ModuliVector() const { return moduli_; }
  const RnsErrorParams<ModularInt>* ErrorParams() const { return error_params_; }
  int PowerOfS() const { return power_of_s_; }
  double Error() const { return error_; }
  const std::vector<RnsPolynomial<ModularInt>>& Components() const { return components_; }

 protected:
  std::vector<RnsPolynomial<ModularInt>>& components() { return components_; }
  ModuliVec& moduli() { return moduli_; }

 private:
  Status CheckAbsorbOperand(const RnsRlweCiphertext& ctxt) const {
    if (Level() != ctxt.Level()) return InvalidArgumentError("`ctxt` has a mismatched level.");
    if (PowerOfS() != ctxt.PowerOfS())
      return InvalidArgumentError("`ctxt` has a mismatched key power.");
    return OkStatus();
  }
  Status CheckAbsorbOperands(const RnsRlweCiphertext& c1, const RnsRlweCiphertext& c2) const {
    if (Level() != c1.Level() || Level() != c2.Level())
      return InvalidArgumentError("`ctxt` has a mismatched level.");
    if (PowerOfS() != c1.PowerOfS() || PowerOfS() != c2.PowerOfS())
      return InvalidArgumentError("`ctxt` has a mismatched key power.");
    return OkStatus();
  }

  std::vector<RnsPolynomial<ModularInt>> components_;
  std::vector<LazyRnsPolynomial<ModularInt>> lazy_components_;
  ModuliVec moduli_;
  const RnsErrorParams<ModularInt>* error_params_;
  int power_of_s_;
  double error_;
};

template <typename ModularInt>
class RnsBgvCiphertext : public RnsRlweCiphertext<ModularInt> {
 public:
  using Base = RnsRlweCiphertext<ModularInt>;
  using Base::Base;
  explicit RnsBgvCiphertext(Base ct) : Base(std::move(ct)) {}

  static StatusOr<RnsBgvCiphertext> Deserialize(const SerializedRnsRlweCiphertext& serialized,
                                                typename Base::ModuliVec moduli,
                                                const RnsErrorParams<ModularInt>* error_params) {
    RLWE_ASSIGN_OR_RETURN(Base ct, Base::Deserialize(serialized, std::move(moduli), error_params));
    return RnsBgvCiphertext(std::move(ct));
  }

  StatusOr<RnsBgvCiphertext> operator*(const RnsBgvCiphertext& that) const {
    RLWE_ASSIGN_OR_RETURN(Base product, this->Mul(that));
    return RnsBgvCiphertext(std::move(product));
  }

  // rns/rns_bgv_ciphertext.h:146-157: one fused ModReduceLsb launch per
  // component.
  Status ModReduce(typename ModularInt::Int t, const RnsInt<ModularInt>& ql_inv) {
    if (this->moduli().size() <= 1)
      return FailedPreconditionError(
          "Cannot perform ModReduce with insufficient number of prime moduli.");
    for (auto& c : this->components())
      RLWE_RETURN_IF_ERROR(c.ModReduceLsb(t, ql_inv, this->Moduli()));
    this->moduli().pop_back();
    return OkStatus();
  }

  // rns/rns_bgv_ciphertext.h:161-174
  StatusOr<RnsBgvCiphertext> Substitute(int substitution_power) const {
    std::vector<RnsPolynomial<ModularInt>> subbed;
    for (const auto& c : this->Components()) {
      RLWE_ASSIGN_OR_RETURN(auto s, c.Substitute(substitution_power, this->Moduli()));
      subbed.push_back(std::move(s));
    }
    int power_of_s = (this->PowerOfS() * substitution_power) % (2 * this->Components()[0].NumCoeffs());
    return RnsBgvCiphertext(std::move(subbed), this->ModuliVector(), power_of_s, this->Error(),
                            this->ErrorParams());
  }
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_CIPHERTEXT_H_
