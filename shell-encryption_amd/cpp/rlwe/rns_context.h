// RnsContext (rns/rns_context.h:46-325): the main modulus chain and the CRT
// constants the hot path consumes (rns/rns_context.cc:169-282).  Auxiliary
// moduli (BFV basis extension) are outside this path.
#ifndef SHG_RLWE_RNS_CONTEXT_H_
#define SHG_RLWE_RNS_CONTEXT_H_

#include <memory>
#include <string>
#include <vector>

#include "rlwe/rns_modulus.h"

namespace rlwe {

template <typename ModularInt>
class RnsContext {
 public:
  using Int = typename ModularInt::Int;

  static StatusOr<RnsContext> CreateForBgv(int log_n, const std::vector<Int>& qs,
                                           Int plaintext_modulus) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (qs.empty()) return InvalidArgumentError("`qs` must not be empty.");
    if (plaintext_modulus < 2) return InvalidArgumentError("`plaintext_modulus` must be > 1.");
    RnsContext ctx;
    ctx.log_n_ = log_n;
    ctx.t_ = plaintext_modulus;
    for (Int q : qs) {
      RLWE_ASSIGN_OR_RETURN(auto pm, PrimeModulus<ModularInt>::Create(q, log_n));
      ctx.owned_.push_back(std::move(pm));
    }
    for (auto& p : ctx.owned_) ctx.moduli_.push_back(p.get());
    RLWE_RETURN_IF_ERROR(ctx.GenerateCrtConstants());
    return ctx;
  }

  int LogN() const { return log_n_; }
  int NumCoeffs() const { return 1 << log_n_; }
  Int PlaintextModulus() const { return t_; }
  std::vector<const PrimeModulus<ModularInt>*> MainPrimeModuli() const { return moduli_; }

  // The i'th entry is q_i^-1 mod q_j for all j (1 at j == i).
  Span<const RnsInt<ModularInt>> MainPrimeModulusInverseResidues() const {
    return Span<const RnsInt<ModularInt>>(prime_q_invs_.data(), prime_q_invs_.size());
  }

  // \hat{Q}_{l,i} = Q_l / q_i mod q_i (rns/rns_context.cc:234-257).
  StatusOr<std::vector<ModularInt>> MainPrimeModulusComplements(int level) const {
    if (level < 0 || level >= (int)moduli_.size())
      return InvalidArgumentError("`level` must be non-negative and at most " +
                                  std::to_string(moduli_.size() - 1));
    std::vector<ModularInt> out;
    for (int i = 0; i <= level; ++i) {
      auto params = moduli_[i]->ModParams();
      auto v = ModularInt::ImportOne(params);
      for (int j = 0; j <= level; ++j)
        if (j != i) v.MulInPlace(prime_qs_[j].zs[i], params);
      out.push_back(v);
    }
    return out;
  }

  // \hat{Q}_{l,i}^-1 mod q_i (rns/rns_context.cc:259-282).
  StatusOr<std::vector<ModularInt>> MainPrimeModulusCrtFactors(int level) const {
    if (level < 0 || level >= (int)moduli_.size())
      return InvalidArgumentError("`level` must be non-negative and at most " +
                                  std::to_string(moduli_.size() - 1));
    std::vector<ModularInt> out;
    for (int i = 0; i <= level; ++i) {
      auto params = moduli_[i]->ModParams();
      auto v = ModularInt::ImportOne(params);
      for (int j = 0; j <= level; ++j)
        if (j != i) v.MulInPlace(prime_q_invs_[j].zs[i], params);
      out.push_back(v);
    }
    return out;
  }

 private:
  // rns/rns_context.cc:169-199
  Status GenerateCrtConstants() {
    for (size_t i = 0; i < moduli_.size(); ++i) {
      RnsInt<ModularInt> qi, qi_inv;
      for (size_t j = 0; j < moduli_.size(); ++j) {
        auto pj = moduli_[j]->ModParams();
        if (i == j) {
          qi.zs.push_back(ModularInt::ImportZero(pj));
          qi_inv.zs.push_back(ModularInt::ImportOne(pj));
        } else {
          RLWE_ASSIGN_OR_RETURN(auto v, ModularInt::ImportInt(moduli_[i]->Modulus(), pj));
          qi_inv.zs.push_back(v.MultiplicativeInverse(pj));
          qi.zs.push_back(v);
        }
      }
      prime_qs_.push_back(std::move(qi));
      prime_q_invs_.push_back(std::move(qi_inv));
    }
    return OkStatus();
  }

  int log_n_ = 0;
  Int t_ = 0;
  std::vector<std::shared_ptr<const PrimeModulus<ModularInt>>> owned_;
  std::vector<const PrimeModulus<ModularInt>*> moduli_;
  std::vector<RnsInt<ModularInt>> prime_qs_;
  std::vector<RnsInt<ModularInt>> prime_q_invs_;
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_CONTEXT_H_
