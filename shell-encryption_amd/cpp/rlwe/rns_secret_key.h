// RnsRlweSecretKey::Sample (rns/rns_secret_key.cc:33-55) and the error
// sampling it and key generation use: SampleFromErrorDistribution
// (sample_error.h:33-83), a centered binomial of the given variance drawn
// from a SecurePrng, and SampleError (rns/error_distribution.h:33-55), which
// lifts those coefficients from the first modulus to every modulus in
// balanced form and transforms them (ConvertBalancedFromPolynomialCoeffs).
// The draws are made on the host in the reference's order; the lift and
// transform run on the device.
#ifndef SHG_RLWE_RNS_SECRET_KEY_H_
#define SHG_RLWE_RNS_SECRET_KEY_H_

#include <string>
#include <vector>

#include "rlwe/prng.h"
#include "rlwe/rns_polynomial.h"

namespace rlwe {

constexpr uint64_t kMaxVariance = 256;  // constants.h:77

// sample_error.h:33-83: for each coefficient, variance * 2 pairs of random
// bits (64 at a time, then 8 at a time, then a masked byte), the popcount
// of the first draw of a pair added and of the second subtracted, starting
// from q and reduced once: a value in [q - 2v, q + 2v) taken mod q.
template <typename ModularInt>
StatusOr<std::vector<ModularInt>> SampleFromErrorDistribution(
    unsigned num_coeffs, uint64_t variance, SecurePrng* prng,
    const typename ModularInt::Params* modulus_params) {
  if (variance > kMaxVariance)
    return InvalidArgumentError("The variance, " + std::to_string(variance) +
                                ", must be at most " + std::to_string(kMaxVariance) + ".");
  std::vector<ModularInt> coeffs;
  coeffs.reserve(num_coeffs);
  const uint64_t q = modulus_params->modulus;
  for (unsigned i = 0; i < num_coeffs; ++i) {
    uint64_t coefficient = q;
    uint64_t k = variance << 1;
    while (k > 0) {
      if (k >= 64) {
        RLWE_ASSIGN_OR_RETURN(uint64_t r, prng->Rand64());
        coefficient += (uint64_t)__builtin_popcountll(r);
        RLWE_ASSIGN_OR_RETURN(r, prng->Rand64());
        coefficient -= (uint64_t)__builtin_popcountll(r);
        k -= 64;
      } else if (k >= 8) {
        RLWE_ASSIGN_OR_RETURN(uint8_t r, prng->Rand8());
        coefficient += (uint64_t)__builtin_popcount(r);
        RLWE_ASSIGN_OR_RETURN(r, prng->Rand8());
        coefficient -= (uint64_t)__builtin_popcount(r);
        k -= 8;
      } else {
        const uint8_t mask = (uint8_t)((1u << k) - 1);
        RLWE_ASSIGN_OR_RETURN(uint8_t r, prng->Rand8());
        coefficient += (uint64_t)__builtin_popcount(r & mask);
        RLWE_ASSIGN_OR_RETURN(r, prng->Rand8());
        coefficient -= (uint64_t)__builtin_popcount(r & mask);
        break;
      }
    }
    if (coefficient >= q) coefficient -= q;
    RLWE_ASSIGN_OR_RETURN(ModularInt c, ModularInt::ImportInt(coefficient, modulus_params));
    coeffs.push_back(c);
  }
  return coeffs;
}

// rns/error_distribution.h:33-55.
template <typename ModularInt>
StatusOr<RnsPolynomial<ModularInt>> SampleError(int log_n, int variance,
                                                 ModuliSpan<ModularInt> moduli,
                                                 SecurePrng* prng) {
  if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
  if (variance <= 0) return InvalidArgumentError("`variance` must be positive.");
  if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
  const auto* p0 = moduli[0]->ModParams();
  RLWE_ASSIGN_OR_RETURN(std::vector<ModularInt> e,
                        SampleFromErrorDistribution<ModularInt>(1u << log_n, (uint64_t)variance,
                                                                prng, p0));
  return RnsPolynomial<ModularInt>::ConvertBalancedFromPolynomialCoeffs(e, p0, moduli);
}

template <typename ModularInt>
class RnsRlweSecretKey {
 public:
  // rns/rns_secret_key.cc:33-55: s drawn from the error distribution, NTT
  // form.
  static StatusOr<RnsRlweSecretKey> Sample(int log_n, int variance,
                                           std::vector<const PrimeModulus<ModularInt>*> moduli,
                                           SecurePrng* prng) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (variance <= 0) return InvalidArgumentError("`variance` must be positive.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must contain at least one element.");
    if (prng == nullptr) return InvalidArgumentError("`prng` must not be null.");
    RLWE_ASSIGN_OR_RETURN(
        RnsPolynomial<ModularInt> s,
        SampleError<ModularInt>(log_n, variance,
                                ModuliSpan<ModularInt>(moduli.data(), moduli.size()), prng));
    return RnsRlweSecretKey(std::move(s), std::move(moduli), variance);
  }

  const RnsPolynomial<ModularInt>& Key() const { return key_; }
  int LogN() const { return key_.LogN(); }
  int NumCoeffs() const { return key_.NumCoeffs(); }
  int Level() const { return (int)moduli_.size() - 1; }
  int Variance() const { return variance_; }
  ModuliSpan<ModularInt> Moduli() const {
    return ModuliSpan<ModularInt>(moduli_.data(), moduli_.size());
  }

 private:
  RnsRlweSecretKey(RnsPolynomial<ModularInt> key,
                   std::vector<const PrimeModulus<ModularInt>*> moduli, int variance)
      : key_(std::move(key)), moduli_(std::move(moduli)), variance_(variance) {}
  RnsPolynomial<ModularInt> key_;
  std::vector<const PrimeModulus<ModularInt>*> moduli_;
  int variance_;
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_SECRET_KEY_H_
