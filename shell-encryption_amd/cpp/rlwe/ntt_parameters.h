// NttParameters and InitializeNttParameters for MontgomeryInt<Uint64>
// (shell_encryption/ntt_parameters.h:41-235, ntt_parameters.cc:21-41):
// psi is the smallest-t primitive 2n-th root; forward twiddles are psi powers
// in bit-reversed order, inverse twiddles psi^-(brv(i)+1), each with its Shoup
// constant.  Built on the host once per (q, n) and uploaded to the device.
#ifndef SHG_RLWE_NTT_PARAMETERS_H_
#define SHG_RLWE_NTT_PARAMETERS_H_

#include <algorithm>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "rlwe/montgomery.h"

namespace rlwe {

constexpr unsigned int kMaxLogNumCoeffs = 15;  // constants.h:75-76

namespace internal {

inline unsigned int Bitrev(unsigned int input, unsigned int log_n) {
  unsigned int output = 0;
  for (unsigned int i = 0; i < log_n; i++) {
    output <<= 1;
    output |= input & 0x01;
    input >>= 1;
  }
  return output;
}

inline std::vector<unsigned int> BitrevArray(unsigned int log_n) {
  unsigned int n = 1u << log_n;
  std::vector<unsigned int> out(n);
  for (unsigned int i = 0; i < n; i++) out[i] = Bitrev(i, log_n);
  return out;
}

template <typename ModularInt>
StatusOr<ModularInt> PrimitiveNthRootOfUnity(unsigned int log_n,
                                             const typename ModularInt::Params* params) {
  typename ModularInt::Int n = params->One() << log_n;
  typename ModularInt::Int half_n = n >> 1;
  typename ModularInt::Int k = (params->modulus - params->One()) / n;
  ModularInt one = ModularInt::ImportOne(params);
  for (typename ModularInt::Int t = params->Two(); t < params->modulus; t++) {
    ModularInt mt = ModularInt::ImportInt(t, params).value();
    ModularInt candidate = mt.ModExp(k, params);
    if (candidate.ModExp(half_n, params) != one) return candidate;
  }
  return UnknownError("Loop in PrimitiveNthRootOfUnity terminated.");
}

// Powers psi^0 .. psi^(n-1), computed by running products (equal to the
// reference's per-entry ModExp, which is exact modular arithmetic).
template <typename ModularInt>
StatusOr<std::vector<ModularInt>> NttPsis(unsigned int log_n,
                                          const typename ModularInt::Params* params) {
  RLWE_ASSIGN_OR_RETURN(ModularInt psi, PrimitiveNthRootOfUnity<ModularInt>(log_n + 1, params));
  unsigned int n = 1u << log_n;
  std::vector<ModularInt> row;
  row.reserve(n);
  ModularInt cur = ModularInt::ImportOne(params);
  for (unsigned int i = 0; i < n; i++) {
    row.push_back(cur);
    cur.MulInPlace(psi, params);
  }
  return row;
}

}  // namespace internal

template <typename ModularInt>
StatusOr<std::vector<ModularInt>> NttPsisBitrev(unsigned int log_n,
                                                const typename ModularInt::Params* params) {
  RLWE_ASSIGN_OR_RETURN(std::vector<ModularInt> psis, internal::NttPsis<ModularInt>(log_n, params));
  std::vector<unsigned int> br = internal::BitrevArray(log_n);
  std::vector<ModularInt> out;
  out.reserve(psis.size());
  for (size_t i = 0; i < psis.size(); i++) out.push_back(psis[br[i]]);
  return out;
}

template <typename ModularInt>
StatusOr<std::vector<ModularInt>> NttPsisInvBitrev(unsigned int log_n,
                                                   const typename ModularInt::Params* params) {
  RLWE_ASSIGN_OR_RETURN(std::vector<ModularInt> row, internal::NttPsis<ModularInt>(log_n, params));
  std::reverse(row.begin() + 1, row.end());
  ModularInt psi_inv = row[row.size() > 1 ? 1 : 0].Negate(params);
  ModularInt negative_psi_inv = row[row.size() > 1 ? 1 : 0];
  std::vector<unsigned int> br = internal::BitrevArray(log_n);
  std::vector<ModularInt> out;
  out.reserve(row.size());
  for (size_t i = 0; i < row.size(); i++) out.push_back(row[br[i]]);
  out[0].MulInPlace(psi_inv, params);
  for (size_t i = 1; i < out.size(); i++) out[i].MulInPlace(negative_psi_inv, params);
  return out;
}

template <typename ModularInt>
struct NttParameters {
  NttParameters() = default;
  NttParameters(const NttParameters&) = delete;
  NttParameters& operator=(const NttParameters&) = delete;
  NttParameters(NttParameters&&) = default;
  NttParameters& operator=(NttParameters&&) = default;

  int number_coeffs = 0;
  std::optional<ModularInt> n_inv_ptr;
  std::optional<ModularInt> psi;  // the primitive 2n-th root used
  std::vector<ModularInt> psis_bitrev;
  std::vector<std::tuple<typename ModularInt::Int, typename ModularInt::Int>> psis_bitrev_constant;
  std::vector<ModularInt> psis_inv_bitrev;
  std::vector<std::tuple<typename ModularInt::Int, typename ModularInt::Int>>
      psis_inv_bitrev_constant;
  std::vector<unsigned int> bitrevs;
};

// ntt_parameters.h:183-235
template <typename ModularInt>
StatusOr<NttParameters<ModularInt>> InitializeNttParameters(
    int log_n, const typename ModularInt::Params* params) {
  if (log_n <= 0) {
    return InvalidArgumentError("log_n must be positive");
  } else if (log_n > static_cast<int>(kMaxLogNumCoeffs)) {
    return InvalidArgumentError("log_n, " + std::to_string(log_n) + ", must be less than " +
                                std::to_string(kMaxLogNumCoeffs) + ".");
  }
  if (!ModularInt::Params::DoesLogNFit(log_n)) {
    return InvalidArgumentError("log_n, " + std::to_string(log_n) +
                                ", does not fit into underlying ModularInt::Int type.");
  }
  NttParameters<ModularInt> output;
  output.number_coeffs = 1 << log_n;
  typename ModularInt::Int two_times_n = params->One() << (log_n + 1);
  if (params->modulus % two_times_n != params->One()) {
    return InvalidArgumentError("modulus is not 1 mod 2n for logn, " + std::to_string(log_n));
  }
  typename ModularInt::Int n = params->One() << log_n;
  RLWE_ASSIGN_OR_RETURN(ModularInt mn, ModularInt::ImportInt(n, params));
  output.n_inv_ptr = mn.MultiplicativeInverse(params);
  RLWE_ASSIGN_OR_RETURN(ModularInt psi,
                        internal::PrimitiveNthRootOfUnity<ModularInt>(log_n + 1, params));
  output.psi = psi;
  RLWE_ASSIGN_OR_RETURN(output.psis_bitrev, NttPsisBitrev<ModularInt>(log_n, params));
  RLWE_ASSIGN_OR_RETURN(output.psis_inv_bitrev, NttPsisInvBitrev<ModularInt>(log_n, params));
  output.psis_bitrev_constant.reserve(output.psis_bitrev.size());
  for (const auto& w : output.psis_bitrev) output.psis_bitrev_constant.push_back(w.GetConstant(params));
  output.psis_inv_bitrev_constant.reserve(output.psis_inv_bitrev.size());
  for (const auto& w : output.psis_inv_bitrev)
    output.psis_inv_bitrev_constant.push_back(w.GetConstant(params));
  output.bitrevs = internal::BitrevArray(log_n);
  return output;
}

}  // namespace rlwe

#endif  // SHG_RLWE_NTT_PARAMETERS_H_
