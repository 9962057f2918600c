// PrimeModulus (rns/rns_modulus.h:30-42) and RnsInt (rns/rns_integer.h:29-52).
// A PrimeModulus owns the host-side MontgomeryIntParams and NttParameters, as
// in the reference, plus the handle of the same prime's device tables.
#ifndef SHG_RLWE_RNS_MODULUS_H_
#define SHG_RLWE_RNS_MODULUS_H_

#include <memory>
#include <vector>

#include "rlwe/device.h"
#include "rlwe/montgomery.h"
#include "rlwe/ntt_parameters.h"

namespace rlwe {

template <typename ModularInt>
struct PrimeModulus {
  using ModularIntParams = typename ModularInt::Params;
  using NttParamsType = NttParameters<ModularInt>;

  std::unique_ptr<const ModularIntParams> mod_params;
  std::unique_ptr<const NttParamsType> ntt_params;
  int device_prime = -1;  // shg_prime_create handle
  int log_n = 0;

  const ModularIntParams* ModParams() const { return mod_params.get(); }
  const NttParamsType* NttParams() const { return ntt_params.get(); }
  typename ModularInt::Int Modulus() const { return mod_params->modulus; }

  // Builds the modulus for ring degree 2^log_n (MontgomeryIntParams::Create
  // and InitializeNttParameters) and registers its tables on the device.
  static StatusOr<std::unique_ptr<const PrimeModulus>> Create(typename ModularInt::Int q,
                                                              int log_n) {
    RLWE_ASSIGN_OR_RETURN(auto mp, ModularIntParams::Create(q));
    RLWE_ASSIGN_OR_RETURN(auto ntt, InitializeNttParameters<ModularInt>(log_n, mp.get()));
    auto pm = std::make_unique<PrimeModulus>();
    pm->mod_params = std::move(mp);
    pm->ntt_params = std::make_unique<const NttParamsType>(std::move(ntt));
    pm->log_n = log_n;
    int id = -1;
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_prime_create(device::Context(), q, log_n, &id)));
    pm->device_prime = id;
    return std::unique_ptr<const PrimeModulus>(std::move(pm));
  }
};

template <typename ModularInt>
struct RnsInt {
  std::vector<ModularInt> zs;
  const std::vector<ModularInt>& RnsRep() const { return zs; }
  int NumModuli() const { return static_cast<int>(zs.size()); }
  const ModularInt& Component(int i) const { return zs[i]; }
  RnsInt Prefix(int len) const {
    return RnsInt{std::vector<ModularInt>(zs.begin(), zs.begin() + len)};
  }
};

template <typename ModularInt>
using ModuliSpan = Span<const PrimeModulus<ModularInt>* const>;

// Device prime handles of a moduli span (at most SHG_MAX_LIMBS).
template <typename ModularInt>
inline std::vector<int> DevicePrimes(ModuliSpan<ModularInt> moduli) {
  std::vector<int> ids;
  ids.reserve(moduli.size());
  for (size_t i = 0; i < moduli.size(); ++i) ids.push_back(moduli[i]->device_prime);
  return ids;
}

// Device prime handles of a moduli span for ring degree 2^log_n: the
// moduli's own handles, or handles registered for that degree.
template <typename ModularInt>
inline StatusOr<std::vector<int>> DevicePrimesAt(ModuliSpan<ModularInt> moduli, int log_n) {
  std::vector<int> ids;
  ids.reserve(moduli.size());
  for (size_t i = 0; i < moduli.size(); ++i) {
    if (moduli[i]->log_n == log_n) {
      ids.push_back(moduli[i]->device_prime);
      continue;
    }
    int id = -1;
    RLWE_RETURN_IF_ERROR(
        device::FromCode(shg_prime_create(device::Context(), moduli[i]->Modulus(), log_n, &id)));
    ids.push_back(id);
  }
  return ids;
}

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_MODULUS_H_
