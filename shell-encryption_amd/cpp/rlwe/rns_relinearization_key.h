// RnsRelinKey (rns/rns_relinearization_key.h:51-184) and RnsGaloisKey
// (rns/rns_galois_key.h:57-310) with device-resident key matrices.
//
// Key generation draws fresh OS-random seeds in the reference, so keys are
// inputs here: CreateFromKeyComponents builds either key from its "a" and "b"
// columns (the reference offers that factory for RnsGaloisKey,
// rns/rns_galois_key.h:122-127).  ApplyTo runs the fused key-switch kernel;
// ApplyToBatch takes many ciphertexts in one launch.
#ifndef SHG_RLWE_RNS_RELINEARIZATION_KEY_H_
#define SHG_RLWE_RNS_RELINEARIZATION_KEY_H_

#include <memory>
#include <string>
#include <vector>

#include "rlwe/rns_ciphertext.h"

namespace rlwe {

namespace internal {

// Key columns packed back to back on the device: polys x L x n words.
template <typename ModularInt>
inline StatusOr<std::shared_ptr<device::Buffer>> PackKeys(
    const std::vector<RnsPolynomial<ModularInt>>& polys) {
  if (polys.empty()) return InvalidArgumentError("key columns must not be empty.");
  const size_t words = (size_t)polys[0].NumModuli() << polys[0].LogN();
  auto buf = std::make_shared<device::Buffer>(words * polys.size());
  for (size_t i = 0; i < polys.size(); ++i) {
    if (!polys[i].IsNttForm()) return InvalidArgumentError("key polynomials must be in NTT form.");
    RLWE_RETURN_IF_ERROR(device::CopyWords(buf->data() + i * words, polys[i].DeviceData(), words));
  }
  return buf;
}

// Packs the components of `cts` (each of degree `degree`) back to back.
template <typename ModularInt>
inline StatusOr<std::shared_ptr<device::Buffer>> PackCiphertexts(
    const std::vector<const RnsRlweCiphertext<ModularInt>*>& cts, int* ntt_mask) {
  const auto& c0 = cts[0]->Components()[0];
  const size_t words = (size_t)c0.NumModuli() << c0.LogN();
  const int len = cts[0]->Len();
  auto buf = std::make_shared<device::Buffer>(words * len * cts.size());
  *ntt_mask = 0;
  for (int i = 0; i < len; ++i)
    if (cts[0]->Components()[i].IsNttForm()) *ntt_mask |= 1 << i;
  for (size_t c = 0; c < cts.size(); ++c) {
    if (cts[c]->Len() != len) return InvalidArgumentError("ciphertexts in a batch must share a degree.");
    for (int i = 0; i < len; ++i) {
      const auto& p = cts[c]->Components()[i];
      if (((*ntt_mask >> i) & 1) != (p.IsNttForm() ? 1 : 0))
        return InvalidArgumentError("ciphertexts in a batch must share component forms.");
      RLWE_RETURN_IF_ERROR(
          device::CopyWords(buf->data() + (c * len + i) * words, p.DeviceData(), words));
    }
  }
  return buf;
}

}  // namespace internal

template <typename ModularInt>
class RnsRelinKey {
 public:
  static StatusOr<RnsRelinKey> CreateFromKeyComponents(
      std::vector<RnsPolynomial<ModularInt>> key_as, std::vector<RnsPolynomial<ModularInt>> key_bs,
      int degree, const RnsGadget<ModularInt>* gadget,
      std::vector<const PrimeModulus<ModularInt>*> moduli) {
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    if (degree < 2) return InvalidArgumentError("`degree` must be at least 2.");
    const size_t need = (size_t)gadget->Dimension() * (degree - 1);
    if (key_as.size() != need || key_bs.size() != need)
      return InvalidArgumentError("`key_as` and `key_bs` must contain " + std::to_string(need) +
                                  " polynomials.");
    RnsRelinKey key;
    RLWE_ASSIGN_OR_RETURN(key.d_as_, internal::PackKeys(key_as));
    RLWE_ASSIGN_OR_RETURN(key.d_bs_, internal::PackKeys(key_bs));
    key.key_as_ = std::move(key_as);
    key.key_bs_ = std::move(key_bs);
    key.gadget_ = gadget;
    key.degree_ = degree;
    key.moduli_ = std::move(moduli);
    return key;
  }

  // rns/rns_relinearization_key.h:82-92
  StatusOr<RnsBgvCiphertext<ModularInt>> ApplyTo(const RnsBgvCiphertext<ModularInt>& ct) const {
    RLWE_ASSIGN_OR_RETURN(auto v, ApplyToBatch({&ct}));
    return std::move(v[0]);
  }

  // Relinearizes all ciphertexts of `cts` in one fused launch.
  StatusOr<std::vector<RnsBgvCiphertext<ModularInt>>> ApplyToBatch(
      const std::vector<const RnsBgvCiphertext<ModularInt>*>& cts) const {
    if (cts.empty()) return std::vector<RnsBgvCiphertext<ModularInt>>{};
    std::vector<const RnsRlweCiphertext<ModularInt>*> base;
    for (auto* c : cts) {
      if (c->Degree() > degree_)
        return InvalidArgumentError("`ciphertext` degree is larger than degree of this "
                                    "relinearization key, " + std::to_string(degree_) + ".");
      if (c->NumModuli() != (int)moduli_.size())
        return InvalidArgumentError("`ciphertext` does not have a matching RNS moduli set.");
      if (c->PowerOfS() != 1)
        return InvalidArgumentError("Relinearization key can only apply to a ciphertext of power 1.");
      base.push_back(c);
    }
    int mask = 0;
    RLWE_ASSIGN_OR_RETURN(auto comps, internal::PackCiphertexts(base, &mask));
    const int L = (int)moduli_.size();
    const int log_n = cts[0]->LogN();
    const size_t words = (size_t)L << log_n;
    device::Buffer o0(words * cts.size()), o1(words * cts.size());
    auto ids = DevicePrimes<ModularInt>(ModuliSpan<ModularInt>(moduli_.data(), moduli_.size()));
    std::vector<int> lbs(gadget_->LogBases().begin(), gadget_->LogBases().end());
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_relin_apply(
        device::Context(), comps->data(), cts[0]->Degree(), mask, d_as_->data(), d_bs_->data(),
        (int)key_as_.size(), lbs.data(), o0.data(), o1.data(), (int)cts.size(), L, ids.data())));
    std::vector<RnsBgvCiphertext<ModularInt>> out;
    for (size_t c = 0; c < cts.size(); ++c) {
      std::vector<RnsPolynomial<ModularInt>> comps_new;
      comps_new.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, o0.data() + c * words));
      comps_new.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, o1.data() + c * words));
      double error = cts[c]->Error() + cts[c]->ErrorParams()->BoundOnGadgetBasedKeySwitching(
                                           cts[c]->Degree() - 1, gadget_->LogGadgetBase(),
                                           gadget_->Dimension());
      out.emplace_back(std::move(comps_new), moduli_, 1, error, cts[c]->ErrorParams());
    }
    return out;
  }

  const std::vector<RnsPolynomial<ModularInt>>& GetKeyA() const { return key_as_; }
  const std::vector<RnsPolynomial<ModularInt>>& GetKeyB() const { return key_bs_; }
  const RnsGadget<ModularInt>* Gadget() const { return gadget_; }
  int Dimension() const { return gadget_->Dimension(); }
  int Degree() const { return degree_; }

 private:
  RnsRelinKey() = default;
  std::vector<RnsPolynomial<ModularInt>> key_as_, key_bs_;
  std::shared_ptr<device::Buffer> d_as_, d_bs_;
  const RnsGadget<ModularInt>* gadget_ = nullptr;
  int degree_ = 0;
  std::vector<const PrimeModulus<ModularInt>*> moduli_;
};

template <typename ModularInt>
class RnsGaloisKey {
 public:
  // rns/rns_galois_key.cc:200-220
  static StatusOr<RnsGaloisKey> CreateFromKeyComponents(
      std::vector<RnsPolynomial<ModularInt>> key_as, std::vector<RnsPolynomial<ModularInt>> key_bs,
      int power, const RnsGadget<ModularInt>* gadget,
      std::vector<const PrimeModulus<ModularInt>*> moduli) {
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    if ((int)key_as.size() != gadget->Dimension() || (int)key_bs.size() != gadget->Dimension())
      return InvalidArgumentError(
          "`key_as` and `key_bs` must have the same length as the gadget dimension.");
    RnsGaloisKey key;
    RLWE_ASSIGN_OR_RETURN(key.d_as_, internal::PackKeys(key_as));
    RLWE_ASSIGN_OR_RETURN(key.d_bs_, internal::PackKeys(key_bs));
    key.key_as_ = std::move(key_as);
    key.key_bs_ = std::move(key_bs);
    key.gadget_ = gadget;
    key.power_ = power;
    key.moduli_ = std::move(moduli);
    return key;
  }

  StatusOr<RnsBgvCiphertext<ModularInt>> ApplyTo(const RnsBgvCiphertext<ModularInt>& ct) const {
    RLWE_ASSIGN_OR_RETURN(auto v, ApplyToBatch({&ct}));
    return std::move(v[0]);
  }

  // rns/rns_galois_key.cc:238-287 over all of `cts` in one fused launch.
  StatusOr<std::vector<RnsBgvCiphertext<ModularInt>>> ApplyToBatch(
      const std::vector<const RnsBgvCiphertext<ModularInt>*>& cts) const {
    if (cts.empty()) return std::vector<RnsBgvCiphertext<ModularInt>>{};
    std::vector<const RnsRlweCiphertext<ModularInt>*> base;
    for (auto* c : cts) {
      if (c->PowerOfS() != power_)
        return InvalidArgumentError("`ciphertext` does not have a matching substitution power.");
      if (c->NumModuli() != (int)moduli_.size())
        return InvalidArgumentError("`ciphertext` does not have a matching RNS moduli set.");
      if (c->Degree() != 1)
        return InvalidArgumentError("Galois key can only apply to a ciphertext of degree 1.");
      if (!c->Components()[0].IsNttForm())
        return InvalidArgumentError("RNS polynomial `this` must be in NTT form.");
      base.push_back(c);
    }
    int mask = 0;
    RLWE_ASSIGN_OR_RETURN(auto comps, internal::PackCiphertexts(base, &mask));
    const int L = (int)moduli_.size();
    const int log_n = cts[0]->LogN();
    const size_t words = (size_t)L << log_n;
    device::Buffer o0(words * cts.size()), o1(words * cts.size());
    auto ids = DevicePrimes<ModularInt>(ModuliSpan<ModularInt>(moduli_.data(), moduli_.size()));
    std::vector<int> lbs(gadget_->LogBases().begin(), gadget_->LogBases().end());
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_galois_apply(
        device::Context(), comps->data(), (mask >> 1) & 1, d_as_->data(), d_bs_->data(),
        lbs.data(), o0.data(), o1.data(), (int)cts.size(), L, ids.data())));
    std::vector<RnsBgvCiphertext<ModularInt>> out;
    for (size_t c = 0; c < cts.size(); ++c) {
      std::vector<RnsPolynomial<ModularInt>> comps_new;
      comps_new.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, o0.data() + c * words));
      comps_new.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, o1.data() + c * words));
      double error = cts[c]->Error() + cts[c]->ErrorParams()->BoundOnGadgetBasedKeySwitching(
                                           1, gadget_->LogGadgetBase(), gadget_->Dimension());
      out.emplace_back(std::move(comps_new), moduli_, 1, error, cts[c]->ErrorParams());
    }
    return out;
  }

  const std::vector<RnsPolynomial<ModularInt>>& GetKeyA() const { return key_as_; }
  const std::vector<RnsPolynomial<ModularInt>>& GetKeyB() const { return key_bs_; }
  int SubstitutionPower() const { return power_; }
  const RnsGadget<ModularInt>* Gadget() const { return gadget_; }

 private:
  RnsGaloisKey() = default;
  std::vector<RnsPolynomial<ModularInt>> key_as_, key_bs_;
  std::shared_ptr<device::Buffer> d_as_, d_bs_;
  const RnsGadget<ModularInt>* gadget_ = nullptr;
  int power_ = 1;
  std::vector<const PrimeModulus<ModularInt>*> moduli_;
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_RELINEARIZATION_KEY_H_
