// LazyRnsPolynomial<MontgomeryInt<Uint64>> (rns/lazy_rns_polynomial.h:37-223,
// rns/lazy_rns_polynomial.cc:24-254) with its sum resident in device memory.
//
// The reference adds the integer products of Montgomery representations
// into 128-bit coefficients and reduces them only when a level counter
// reaches 2^(127 - 2 log q) (Refresh), then Export reduces once more and
// takes the sum out of Montgomery form: the exported representation is
// sum_t a_t b_t 2^-64 mod q.  Here every update adds its Montgomery product
// to a canonical device accumulator (one launch per update), which holds
// that same residue at all times, so Export is a copy and no refresh is
// needed.  Checks and error messages follow the reference.
#ifndef SHG_RLWE_LAZY_RNS_POLYNOMIAL_H_
#define SHG_RLWE_LAZY_RNS_POLYNOMIAL_H_

#include <string>
#include <utility>

#include "rlwe/rns_polynomial.h"

namespace rlwe {

template <typename ModularInt>
class LazyRnsPolynomial {
 public:
  using Integer = typename ModularInt::Int;
  using Moduli = ModuliSpan<ModularInt>;

  // rns/lazy_rns_polynomial.h:44-79: the product a * b.
  static StatusOr<LazyRnsPolynomial> Create(const RnsPolynomial<ModularInt>& a,
                                            const RnsPolynomial<ModularInt>& b, Moduli moduli) {
    if (!a.IsNttForm() || !b.IsNttForm())
      return InvalidArgumentError("Polynomials `a` and `b` must be in the NTT form.");
    if (b.NumCoeffs() != a.NumCoeffs())
      return InvalidArgumentError(
          "Polynomials `a` and `b` must have the same number of coefficients.");
    const int num_moduli = (int)moduli.size();
    if (a.NumModuli() != num_moduli || b.NumModuli() != num_moduli)
      return InvalidArgumentError("Polynomials `a` and `b` must be defined wrt `moduli`");
    RLWE_RETURN_IF_ERROR(CheckLogMaximalLevel(moduli));
    RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> sum, a.Mul(b, moduli));
    return LazyRnsPolynomial(std::move(sum));
  }

  // rns/lazy_rns_polynomial.h:82-101.
  static StatusOr<LazyRnsPolynomial> CreateZero(int log_n, Moduli moduli) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
    RLWE_RETURN_IF_ERROR(CheckLogMaximalLevel(moduli));
    RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> zero,
                          RnsPolynomial<ModularInt>::CreateZero(log_n, moduli, /*is_ntt=*/true));
    return LazyRnsPolynomial(std::move(zero));
  }

  // rns/lazy_rns_polynomial.h:103-121: the sum reduced wrt `moduli`, NTT
  // form.
  StatusOr<RnsPolynomial<ModularInt>> Export(Moduli moduli) {
    if ((int)moduli.size() != sum_.NumModuli())
      return InvalidArgumentError("`moduli` does not contain enough RNS moduli.");
    return sum_;
  }

  // rns/lazy_rns_polynomial.cc:20-41, 89-107: adds a * b.
  Status FusedMulAddInPlace(const RnsPolynomial<ModularInt>& a,
                            const RnsPolynomial<ModularInt>& b, Moduli moduli) {
    if (!a.IsNttForm() || !b.IsNttForm())
      return InvalidArgumentError("Polynomials `a` and `b` must be in NTT form.");
    const int num_moduli = (int)moduli.size();
    if (a.NumModuli() != num_moduli || b.NumModuli() != num_moduli ||
        sum_.NumModuli() != num_moduli)
      return InvalidArgumentError(
          "Polynomials `a`, `b`, and this must all be defined wrt `moduli`");
    if (a.NumCoeffs() != sum_.NumCoeffs() || b.NumCoeffs() != sum_.NumCoeffs())
      return InvalidArgumentError(
          "Polynomials `a` and `b` must have the same number of coefficients as this lazy "
          "polynomial.");
    return sum_.FusedMulAddInPlace(a, b, moduli);
  }

  // rns/lazy_rns_polynomial.cc:170-188: adds (a + b) * c.
  Status FusedMulSumAddInPlace(const RnsPolynomial<ModularInt>& a,
                               const RnsPolynomial<ModularInt>& b,
                               const RnsPolynomial<ModularInt>& c, Moduli moduli) {
    return SumOrDifference(false, a, b, c, moduli);
  }

  // rns/lazy_rns_polynomial.cc:228-254: adds (a - b) * c.
  Status FusedMulDifferenceAddInPlace(const RnsPolynomial<ModularInt>& a,
                                      const RnsPolynomial<ModularInt>& b,
                                      const RnsPolynomial<ModularInt>& c, Moduli moduli) {
    return SumOrDifference(true, a, b, c, moduli);
  }

 private:
  explicit LazyRnsPolynomial(RnsPolynomial<ModularInt> sum) : sum_(std::move(sum)) {}

  // rns/lazy_rns_polynomial.h:123-150 (ComputeLogMaximalLevel): a 128-bit
  // sum must leave at least one bit of headroom above the product of two
  // residues.
  static Status CheckLogMaximalLevel(Moduli moduli) {
    for (size_t i = 0; i < moduli.size(); ++i) {
      const auto* mp = moduli[i]->ModParams();
      const int level = (int)mp->bitsize_bigint - 1 - 2 * (int)mp->log_modulus;
      if (level < 1)
        return InvalidArgumentError(
            "The RNS moduli do not allow for lazy polynomials, as the logarithm of the maximum "
            "level, " + std::to_string(level) + " is stricly smaller than 1.");
    }
    return OkStatus();
  }

  // rns/lazy_rns_polynomial.cc:58-87 (CheckFusedMulSumAddInPlaceParameters),
  // then one launch of the sum or difference update.
  Status SumOrDifference(bool subtract, const RnsPolynomial<ModularInt>& a,
                         const RnsPolynomial<ModularInt>& b, const RnsPolynomial<ModularInt>& c,
                         Moduli moduli) {
    if (!a.IsNttForm() || !b.IsNttForm() || !c.IsNttForm())
      return InvalidArgumentError("Polynomials `a`, `b`, and `c` must be in NTT form.");
    const int num_moduli = (int)moduli.size();
    if (a.NumModuli() != num_moduli || b.NumModuli() != num_moduli ||
        c.NumModuli() != num_moduli || sum_.NumModuli() != num_moduli)
      return InvalidArgumentError(
          "Polynomials `a`, `b`, `c`, and this must all be defined wrt `moduli`");
    for (int i = 0; i < num_moduli; ++i)
      if (moduli[i]->ModParams()->log_modulus + 3 >= sizeof(Integer) * 8)
        return InvalidArgumentError("Modulus is too large to perform fused multiply-sum-add.");
    const int n = sum_.NumCoeffs();
    if (a.NumCoeffs() != n || b.NumCoeffs() != n || c.NumCoeffs() != n)
      return InvalidArgumentError(
          "Polynomials `a`, `b`, and `c` must have the same number of coefficients as this lazy "
          "polynomial.");
    auto ids = DevicePrimes<ModularInt>(moduli);
    return device::FromCode(shg_fused_mul_sum_add(device::Context(), subtract ? 1 : 0,
                                                  sum_.DeviceData(), a.DeviceData(),
                                                  b.DeviceData(), c.DeviceData(), 1, num_moduli,
                                                  ids.data()));
  }

  RnsPolynomial<ModularInt> sum_;  // canonical Montgomery representations
};

}  // namespace rlwe

#endif  // SHG_RLWE_LAZY_RNS_POLYNOMIAL_H_
