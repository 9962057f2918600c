// RnsGadget (rns/rns_gadget.h:50-117, rns/rns_gadget.cc:42-153) and
// RnsErrorParams' key-switching bound (rns/rns_error_params.h:84-90).
#ifndef SHG_RLWE_RNS_GADGET_H_
#define SHG_RLWE_RNS_GADGET_H_

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "rlwe/rns_polynomial.h"

namespace rlwe {

template <typename ModularInt>
inline int GadgetSize(int log_base, const typename ModularInt::Params* mod_params) {
  return (mod_params->log_modulus + (log_base - 1)) / log_base;  // gadget.h:43-47
}

template <typename ModularInt>
class RnsGadget {
 public:
  using Moduli = ModuliSpan<ModularInt>;

  static StatusOr<RnsGadget> Create(int log_n, std::vector<size_t> log_bs,
                                    Span<const ModularInt> q_hats,
                                    Span<const ModularInt> q_hat_invs, Moduli moduli) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
    const size_t L = moduli.size();
    if (log_bs.size() != L)
      return InvalidArgumentError("`log_bs` must contain " + std::to_string(L) + " elements.");
    if (q_hats.size() != L)
      return InvalidArgumentError("`q_hats` must contain " + std::to_string(L) + " elements.");
    if (q_hat_invs.size() != L)
      return InvalidArgumentError("`q_hat_invs` must contain " + std::to_string(L) + " elements.");
    std::vector<size_t> dims(L);
    std::vector<RnsInt<ModularInt>> gs;
    for (size_t i = 0; i < L; ++i) {
      auto p = moduli[i]->ModParams();
      dims[i] = GadgetSize<ModularInt>((int)log_bs[i], p);
      RLWE_ASSIGN_OR_RETURN(auto bi, ModularInt::ImportInt(uint64_t{1} << log_bs[i], p));
      ModularInt cur = q_hats[i].Mul(q_hat_invs[i], p);
      for (size_t j = 0; j < dims[i]; ++j) {
        RnsInt<ModularInt> e;
        for (size_t l = 0; l < L; ++l) e.zs.push_back(ModularInt::ImportZero(moduli[l]->ModParams()));
        e.zs[i] = cur;
        gs.push_back(std::move(e));
        cur.MulInPlace(bi, p);
      }
    }
    return RnsGadget(std::move(log_bs), std::move(dims), std::move(gs));
  }

  // g^-1(a) on the device: k coefficient-form RNS polynomials.
  StatusOr<std::vector<RnsPolynomial<ModularInt>>> Decompose(const RnsPolynomial<ModularInt>& a,
                                                             Moduli moduli) const {
    if (a.IsNttForm()) return InvalidArgumentError("`a` must be in coefficient form");
    if (a.NumModuli() != (int)moduli.size())
      return InvalidArgumentError("`a` must be defined with respect to " +
                                  std::to_string(moduli.size()) + " RNS moduli.");
    const int k = Dimension();
    const int L = a.NumModuli();
    const size_t words = (size_t)L << a.LogN();
    device::Buffer digits(words * k);
    auto ids = DevicePrimes<ModularInt>(moduli);
    std::vector<int> lbs(log_bs_.begin(), log_bs_.end());
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_gadget_decompose(
        device::Context(), a.DeviceData(), lbs.data(), digits.data(), 1, L, ids.data())));
    std::vector<RnsPolynomial<ModularInt>> out;
    for (int j = 0; j < k; ++j)
      out.push_back(RnsPolynomial<ModularInt>::FromDevice(a.LogN(), L, false,
                                                          digits.data() + j * words));
    return out;
  }

  int LogGadgetBase() const { return (int)*std::max_element(log_bs_.begin(), log_bs_.end()); }
  int Dimension() const { return (int)std::accumulate(dims_.begin(), dims_.end(), size_t{0}); }
  const RnsInt<ModularInt>& Component(int index) const { return gs_[index]; }
  const std::vector<size_t>& LogBases() const { return log_bs_; }

 private:
  RnsGadget(std::vector<size_t> log_bs, std::vector<size_t> dims,
            std::vector<RnsInt<ModularInt>> gs)
      : log_bs_(std::move(log_bs)), dims_(std::move(dims)), gs_(std::move(gs)) {}
  std::vector<size_t> log_bs_;
  std::vector<size_t> dims_;
  std::vector<RnsInt<ModularInt>> gs_;
};

// The error bookkeeping the key-switching classes return with their results
// (rns/rns_error_params.h:84-90, 165-166).
template <typename ModularInt>
class RnsErrorParams {
 public:
  // rns/rns_error_params.h:43-72.  The auxiliary moduli (BFV basis
  // extension) only enter bounds this path does not use.
  static StatusOr<RnsErrorParams> Create(int log_n, ModuliSpan<ModularInt> main_moduli,
                                         ModuliSpan<ModularInt> aux_moduli,
                                         int log_plaintext_modulus, double sigma) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (main_moduli.empty()) return InvalidArgumentError("`main_moduli` must not be empty.");
    if (log_plaintext_modulus <= 0)
      return InvalidArgumentError("`log_plaintext_modulus` must be positive.");
    if (sigma <= 0) return InvalidArgumentError("`sigma` must be positive.");
    (void)aux_moduli;
    return RnsErrorParams(log_n, log_plaintext_modulus, sigma);
  }
  double BoundOnGadgetBasedKeySwitching(int num_components, int log_gadget_base,
                                        int num_digits) const {
    int gadget_base = 1 << log_gadget_base;
    int dimension = 1 << log_n_;
    return (8.0 / std::sqrt(3.0)) * ExportDoubleT() * num_digits * sigma_ * dimension *
           gadget_base * num_components;
  }
  double B_secretkey_encryption() const {
    return ExportDoubleT() * std::sqrt(1 << log_n_) * (std::sqrt(3.0) + 6.0 * sigma_);
  }
  // rns/rns_error_params.h:84-87, 129-172: the size of an NTT coefficient
  // of a random plaintext (the factor an absorb multiplies the error by), of
  // a public-key encryption, and of the rounding term of a modulus switch.
  double B_plaintext() const { return ExportDoubleT() * std::sqrt(3.0 * (1 << log_n_)); }
  double B_publickey_encryption() const {
    const double dimension = 1 << log_n_;
    return ExportDoubleT() * (std::sqrt(dimension) * (6 * sigma_ + std::sqrt(3.0)) +
                              72 * dimension * sigma_ * sigma_);
  }
  double B_scale() const {
    const double dimension = 1 << log_n_;
    return ExportDoubleT() *
           (std::sqrt(3.0 * dimension) + 8.0 * dimension * std::sqrt(1 / 3.0));
  }
  // As in the reference: the square of ceil(log2 t).
  double ExportDoubleT() const { return std::pow(log_plaintext_modulus_, 2); }

 private:
  RnsErrorParams(int log_n, int log_t, double sigma)
      : log_n_(log_n), log_plaintext_modulus_(log_t), sigma_(sigma) {}
  int log_n_;
  int log_plaintext_modulus_;
  double sigma_;
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_GADGET_H_
