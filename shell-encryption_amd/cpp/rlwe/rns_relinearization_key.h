// RnsRelinKey (rns/rns_relinearization_key.h:51-184) and RnsGaloisKey
// (rns/rns_galois_key.h:57-310) with device-resident key matrices.
//
// Key generation follows the reference (rns/rns_relinearization_key.cc:
// 46-182, rns/rns_galois_key.cc:39-199): the pads "a" come from a PRNG seeded
// with a fresh seed (for ChaCha generated on the device by
// shg_sample_uniform_chacha, for HKDF drawn on the host), the errors from a
// second, fresh PRNG, and b = t*e + g_j*s' - a*s is formed with device
// pointwise operations.  CreateFromKeyComponents builds either key from given
// columns (the reference offers it for RnsGaloisKey; for RnsRelinKey it is
// an addition to the reference's API, for keys made elsewhere).  ApplyTo runs the
// fused key-switch kernel; ApplyToBatch takes many ciphertexts in one launch.
// RnsGaloisKey::Compose is one batched Galois key switch of the other key's
// substituted columns (rns/rns_galois_key.cc:349-406).
#ifndef SHG_RLWE_RNS_RELINEARIZATION_KEY_H_
#define SHG_RLWE_RNS_RELINEARIZATION_KEY_H_

#include <memory>
#include <string>
#include <vector>

#include "rlwe/prng.h"
#include "rlwe/rns_ciphertext.h"
#include "rlwe/rns_secret_key.h"
#include "rlwe/serialization.h"

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

// The packed key columns prepared for the key-switch kernels
// (shg_ks_key_create: FP64-form copies when every modulus is below 2^50);
// shared by copies of the key object, released with the last.
template <typename ModularInt>
inline StatusOr<std::shared_ptr<shg_ks_key>> PrepareKey(
    const device::Buffer& as, const device::Buffer& bs, int num_key_polys,
    ModuliSpan<ModularInt> moduli) {
  auto ids = DevicePrimes<ModularInt>(moduli);
  shg_ks_key* k = nullptr;
  RLWE_RETURN_IF_ERROR(device::FromCode(shg_ks_key_create(
      device::Context(), as.data(), bs.data(), num_key_polys, (int)ids.size(), ids.data(), &k)));
  return std::shared_ptr<shg_ks_key>(k, [](shg_ks_key* p) {
    (void)shg_ks_key_destroy(device::Context(), p);
  });
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

// Fresh seed for a PRNG of `prng_type` (GenerateSeed of either PRNG).
inline StatusOr<std::string> GenerateSeedFor(PrngType prng_type) {
  if (prng_type == PRNG_TYPE_HKDF) return SingleThreadHkdfPrng::GenerateSeed();
  if (prng_type == PRNG_TYPE_CHACHA) return SingleThreadChaChaPrng::GenerateSeed();
  return InvalidArgumentError("PrngType not specified correctly.");
}

inline StatusOr<std::unique_ptr<SecurePrng>> CreatePrng(const std::string& seed,
                                                        PrngType prng_type) {
  if (prng_type == PRNG_TYPE_HKDF) {
    RLWE_ASSIGN_OR_RETURN(auto p, SingleThreadHkdfPrng::Create(seed));
    return std::unique_ptr<SecurePrng>(std::move(p));
  }
  if (prng_type == PRNG_TYPE_CHACHA) {
    RLWE_ASSIGN_OR_RETURN(auto p, SingleThreadChaChaPrng::Create(seed));
    return std::unique_ptr<SecurePrng>(std::move(p));
  }
  return InvalidArgumentError("PrngType not specified correctly.");
}

// `count` RnsPolynomial::SampleUniform draws from one PRNG seeded with
// `seed`: on the device for ChaCha, on the host for HKDF.
template <typename ModularInt>
StatusOr<std::vector<RnsPolynomial<ModularInt>>> SamplePads(int count, int log_n,
                                                            ModuliSpan<ModularInt> moduli,
                                                            const std::string& seed,
                                                            PrngType prng_type) {
  std::vector<RnsPolynomial<ModularInt>> pads;
  if (prng_type == PRNG_TYPE_CHACHA) {
    if ((int)seed.size() != SingleThreadChaChaPrng::SeedLength())
      return SingleThreadChaChaPrng::Create(seed).status();
    const int L = (int)moduli.size();
    const size_t words = (size_t)L << log_n;
    device::Buffer buf(words * (size_t)count);
    RLWE_ASSIGN_OR_RETURN(std::vector<int> ids, DevicePrimesAt(moduli, log_n));
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_sample_uniform_chacha(
        device::Context(), reinterpret_cast<const uint8_t*>(seed.data()), seed.size(), count, L,
        ids.data(), buf.data())));
    for (int i = 0; i < count; ++i)
      pads.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, buf.data() + i * words));
    return pads;
  }
  RLWE_ASSIGN_OR_RETURN(std::unique_ptr<SecurePrng> prng, CreatePrng(seed, prng_type));
  for (int i = 0; i < count; ++i) {
    RLWE_ASSIGN_OR_RETURN(auto a, RnsPolynomial<ModularInt>::SampleUniform(log_n, prng.get(), moduli));
    pads.push_back(std::move(a));
  }
  return pads;
}

// out = sum_j terms[j] (.) key[j], Montgomery products, with the lazy
// accumulation of LazyRnsPolynomial (rns/lazy_rns_polynomial.h:84-129):
// terms packed on the device, key columns already packed in `key`.
template <typename ModularInt>
StatusOr<RnsPolynomial<ModularInt>> LazyInnerProduct(
    const std::vector<RnsPolynomial<ModularInt>>& terms, const device::Buffer& key,
    ModuliSpan<ModularInt> moduli) {
  if (terms.empty()) return InvalidArgumentError("no terms.");
  const int L = (int)moduli.size();
  const int log_n = terms[0].LogN();
  const size_t words = (size_t)L << log_n;
  device::Buffer packed(words * terms.size());
  for (size_t j = 0; j < terms.size(); ++j) {
    if (!terms[j].IsNttForm()) return InvalidArgumentError("RNS polynomial `a` must be in NTT form.");
    RLWE_RETURN_IF_ERROR(device::CopyWords(packed.data() + j * words, terms[j].DeviceData(), words));
  }
  device::Buffer out(words);
  auto ids = DevicePrimes<ModularInt>(moduli);
  RLWE_RETURN_IF_ERROR(device::FromCode(shg_lazy_fma(device::Context(), packed.data(), words,
                                                     key.data(), words, (int)terms.size(),
                                                     out.data(), L, ids.data())));
  return RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, out.data());
}

// One key-switching column b = t * e + target - pad * s (NTT form), e drawn
// from `prng` (rns/rns_relinearization_key.cc:160-170,
// rns/rns_galois_key.cc:178-191).
template <typename ModularInt>
StatusOr<RnsPolynomial<ModularInt>> KeyColumn(const RnsPolynomial<ModularInt>& pad,
                                              const RnsPolynomial<ModularInt>& target,
                                              const RnsPolynomial<ModularInt>& s, int variance,
                                              typename ModularInt::Int error_scalar,
                                              SecurePrng* prng, ModuliSpan<ModularInt> moduli) {
  RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> u, pad.Negate(moduli));
  RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> b,
                        SampleError<ModularInt>(s.LogN(), variance, moduli, prng));
  RLWE_RETURN_IF_ERROR(b.MulInPlace(error_scalar, moduli));
  RLWE_RETURN_IF_ERROR(b.AddInPlace(target, moduli));
  RLWE_RETURN_IF_ERROR(b.FusedMulAddInPlace(u, s, moduli));
  return b;
}

}  // namespace internal

template <typename ModularInt>
class RnsRelinKey {
 public:
  using Integer = typename ModularInt::Int;

  // rns/rns_relinearization_key.h:59-73
  static StatusOr<RnsRelinKey> CreateForBgv(const RnsRlweSecretKey<ModularInt>& secret_key,
                                            int degree, int variance,
                                            const RnsGadget<ModularInt>* gadget,
                                            Integer plaintext_modulus, PrngType prng_type) {
    return Create(secret_key, degree, variance, gadget, prng_type, plaintext_modulus);
  }
  static StatusOr<RnsRelinKey> CreateForBfv(const RnsRlweSecretKey<ModularInt>& secret_key,
                                            int degree, int variance,
                                            const RnsGadget<ModularInt>* gadget,
                                            PrngType prng_type) {
    return Create(secret_key, degree, variance, gadget, prng_type);
  }

  // rns/rns_relinearization_key.cc:46-77
  static StatusOr<RnsRelinKey> Create(const RnsRlweSecretKey<ModularInt>& secret_key, int degree,
                                      int variance, const RnsGadget<ModularInt>* gadget,
                                      PrngType prng_type, Integer error_scalar = 1) {
    if (degree <= 1) return InvalidArgumentError("`degree` must be at least 2.");
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    RLWE_ASSIGN_OR_RETURN(std::string seed, internal::GenerateSeedFor(prng_type));
    RLWE_ASSIGN_OR_RETURN(auto pads, SampleRandomPad(gadget->Dimension(), degree,
                                                     secret_key.LogN(), secret_key.Moduli(),
                                                     seed, prng_type));
    return CreateWithRandomPad(std::move(pads), secret_key, degree, variance, gadget, seed,
                               prng_type, error_scalar);
  }

  // rns/rns_relinearization_key.cc:79-111
  static StatusOr<std::vector<RnsPolynomial<ModularInt>>> SampleRandomPad(
      int dimension, int degree, int log_n, ModuliSpan<ModularInt> moduli,
      const std::string& prng_seed, PrngType prng_type) {
    if (dimension <= 0) return InvalidArgumentError("`dimension` must be positive.");
    if (degree <= 1) return InvalidArgumentError("`degree` must be at least 2.");
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (prng_type != PRNG_TYPE_HKDF && prng_type != PRNG_TYPE_CHACHA)
      return InvalidArgumentError("PrngType not specified correctly.");
    return internal::SamplePads<ModularInt>(dimension * (degree - 1), log_n, moduli, prng_seed,
                                            prng_type);
  }

  // rns/rns_relinearization_key.cc:113-182: b = t*e + g_j*s^i - a*s for
  // i = 2..degree, j < k, the errors from a fresh PRNG of the same type.
  static StatusOr<RnsRelinKey> CreateWithRandomPad(
      std::vector<RnsPolynomial<ModularInt>> pads,
      const RnsRlweSecretKey<ModularInt>& secret_key, int degree, int variance,
      const RnsGadget<ModularInt>* gadget, const std::string& prng_pad_seed,
      PrngType prng_type, Integer error_scalar = 1) {
    if (variance <= 0) return InvalidArgumentError("`variance` must be positive.");
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    if (degree <= 1) return InvalidArgumentError("`degree` must be at least 2.");
    const int k = gadget->Dimension();
    if ((int)pads.size() != k * (degree - 1))
      return InvalidArgumentError("`pads` must contain " + std::to_string(k * (degree - 1)) +
                                  " polynomials.");
    // The errors' PRNG: HKDF for HKDF pads, ChaCha otherwise.
    const PrngType enc_type = prng_type == PRNG_TYPE_HKDF ? PRNG_TYPE_HKDF : PRNG_TYPE_CHACHA;
    RLWE_ASSIGN_OR_RETURN(std::string enc_seed, internal::GenerateSeedFor(enc_type));
    RLWE_ASSIGN_OR_RETURN(auto prng_e, internal::CreatePrng(enc_seed, enc_type));
    const RnsPolynomial<ModularInt>& s = secret_key.Key();
    auto moduli = secret_key.Moduli();
    RnsPolynomial<ModularInt> secret = s;
    std::vector<RnsPolynomial<ModularInt>> key_bs;
    int index = 0;
    for (int i = 2; i <= degree; ++i) {
      RLWE_RETURN_IF_ERROR(secret.MulInPlace(s, moduli));  // s^i
      for (int j = 0; j < k; ++j) {
        RnsPolynomial<ModularInt> z = secret;
        RLWE_RETURN_IF_ERROR(z.MulInPlace(gadget->Component(j), moduli));
        RLWE_ASSIGN_OR_RETURN(auto b, internal::KeyColumn(pads[index++], z, s, variance,
                                                          error_scalar, prng_e.get(), moduli));
        key_bs.push_back(std::move(b));
      }
    }
    std::vector<const PrimeModulus<ModularInt>*> mv(moduli.begin(), moduli.end());
    RLWE_ASSIGN_OR_RETURN(RnsRelinKey key, CreateFromKeyComponents(std::move(pads), std::move(key_bs),
                                                                   degree, gadget, std::move(mv)));
    key.prng_seed_ = prng_pad_seed;
    key.prng_type_ = prng_type;
    return key;
  }

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
    RLWE_ASSIGN_OR_RETURN(key.ks_, internal::PrepareKey<ModularInt>(
                                       *key.d_as_, *key.d_bs_, (int)key_as.size(),
                                       ModuliSpan<ModularInt>(moduli.data(), moduli.size())));
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
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_relin_apply_key(
        device::Context(), comps->data(), cts[0]->Degree(), mask, ks_.get(), lbs.data(),
        o0.data(), o1.data(), (int)cts.size(), L, ids.data())));
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
  const std::string& PrngSeed() const { return prng_seed_; }
  PrngType GetPrngType() const { return prng_type_; }

 private:
  RnsRelinKey() = default;
  std::vector<RnsPolynomial<ModularInt>> key_as_, key_bs_;
  std::shared_ptr<device::Buffer> d_as_, d_bs_;
  std::shared_ptr<shg_ks_key> ks_;  // over d_as_, d_bs_ (declared after them)
  const RnsGadget<ModularInt>* gadget_ = nullptr;
  int degree_ = 0;
  std::vector<const PrimeModulus<ModularInt>*> moduli_;
  std::string prng_seed_;
  PrngType prng_type_ = PRNG_TYPE_INVALID;
};

template <typename ModularInt>
class RnsGaloisKey {
 public:
  using Integer = typename ModularInt::Int;
  using ModuliVec = std::vector<const PrimeModulus<ModularInt>*>;

  // rns/rns_galois_key.h:65-76
  static StatusOr<RnsGaloisKey> CreateForBgv(const RnsRlweSecretKey<ModularInt>& secret_key,
                                             int power, int variance,
                                             const RnsGadget<ModularInt>* gadget,
                                             Integer plaintext_modulus, PrngType prng_type) {
    return Create(secret_key, power, variance, gadget, prng_type, plaintext_modulus);
  }
  static StatusOr<RnsGaloisKey> CreateForBfv(const RnsRlweSecretKey<ModularInt>& secret_key,
                                             int power, int variance,
                                             const RnsGadget<ModularInt>* gadget,
                                             PrngType prng_type) {
    return Create(secret_key, power, variance, gadget, prng_type);
  }

  // rns/rns_galois_key.cc:39-79
  static StatusOr<RnsGaloisKey> Create(const RnsRlweSecretKey<ModularInt>& secret_key, int power,
                                       int variance, const RnsGadget<ModularInt>* gadget,
                                       PrngType prng_type, Integer error_scalar = 1) {
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    RLWE_ASSIGN_OR_RETURN(std::string seed, internal::GenerateSeedFor(prng_type));
    RLWE_ASSIGN_OR_RETURN(auto pads, SampleRandomPad(gadget->Dimension(), secret_key.LogN(),
                                                     secret_key.Moduli(), seed, prng_type));
    return CreateWithRandomPad(std::move(pads), secret_key, power, variance, gadget, seed,
                               prng_type, error_scalar);
  }

  // rns/rns_galois_key.cc:81-105: the "a" columns from a PRNG seeded with
  // `prng_seed` (ChaCha pads generated on the device).
  static StatusOr<std::vector<RnsPolynomial<ModularInt>>> SampleRandomPad(
      int dimension, int log_n, ModuliSpan<ModularInt> moduli, const std::string& prng_seed,
      PrngType prng_type) {
    if (prng_type != PRNG_TYPE_HKDF && prng_type != PRNG_TYPE_CHACHA)
      return InvalidArgumentError("PrngType not specified correctly.");
    return internal::SamplePads<ModularInt>(dimension, log_n, moduli, prng_seed, prng_type);
  }

  static StatusOr<RnsGaloisKey> CreateWithRandomPadForBgv(
      std::vector<RnsPolynomial<ModularInt>> key_pads,
      const RnsRlweSecretKey<ModularInt>& secret_key, int power, int variance,
      const RnsGadget<ModularInt>* gadget, Integer plaintext_modulus,
      const std::string& prng_pad_seed, PrngType prng_type) {
    return CreateWithRandomPad(std::move(key_pads), secret_key, power, variance, gadget,
                               prng_pad_seed, prng_type, plaintext_modulus);
  }

  // rns/rns_galois_key.cc:107-199: b_j = t*e + g_j*s(X^power) - a_j*s.
  static StatusOr<RnsGaloisKey> CreateWithRandomPad(
      std::vector<RnsPolynomial<ModularInt>> pads,
      const RnsRlweSecretKey<ModularInt>& secret_key, int power, int variance,
      const RnsGadget<ModularInt>* gadget, const std::string& prng_pad_seed,
      PrngType prng_type, Integer error_scalar = 1) {
    const int num_coeffs = secret_key.NumCoeffs();
    if (power < 0 || (power % 2) == 0 || power >= 2 * num_coeffs)
      return InvalidArgumentError("`power` must be a non-negative odd integer less than 2*n.");
    if (variance <= 0) return InvalidArgumentError("`variance` must be positive.");
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    if (prng_type != PRNG_TYPE_HKDF && prng_type != PRNG_TYPE_CHACHA)
      return InvalidArgumentError("PrngType not specified correctly.");
    const int k = gadget->Dimension();
    if ((int)pads.size() != k)
      return InvalidArgumentError("`pads` must contain " + std::to_string(k) + " polynomials.");
    RLWE_ASSIGN_OR_RETURN(std::string enc_seed, internal::GenerateSeedFor(prng_type));
    RLWE_ASSIGN_OR_RETURN(auto prng_e, internal::CreatePrng(enc_seed, prng_type));
    auto moduli = secret_key.Moduli();
    const RnsPolynomial<ModularInt>& s = secret_key.Key();
    RLWE_ASSIGN_OR_RETURN(RnsPolynomial<ModularInt> source, s.Substitute(power, moduli));
    std::vector<RnsPolynomial<ModularInt>> key_bs;
    for (int i = 0; i < k; ++i) {
      RnsPolynomial<ModularInt> z = source;
      RLWE_RETURN_IF_ERROR(z.MulInPlace(gadget->Component(i), moduli));
      RLWE_ASSIGN_OR_RETURN(auto b, internal::KeyColumn(pads[i], z, s, variance, error_scalar,
                                                        prng_e.get(), moduli));
      key_bs.push_back(std::move(b));
    }
    return CreateFromKeyComponents(std::move(pads), std::move(key_bs), power, gadget,
                                   ModuliVec(moduli.begin(), moduli.end()), prng_pad_seed,
                                   prng_type);
  }

  // rns/rns_galois_key.cc:200-220 (the seed and PRNG type of the "a"
  // columns are kept for Serialize).
  static StatusOr<RnsGaloisKey> CreateFromKeyComponents(
      std::vector<RnsPolynomial<ModularInt>> key_as, std::vector<RnsPolynomial<ModularInt>> key_bs,
      int power, const RnsGadget<ModularInt>* gadget, ModuliVec moduli,
      const std::string& prng_seed = std::string(), PrngType prng_type = PRNG_TYPE_INVALID) {
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    if ((int)key_as.size() != gadget->Dimension() || (int)key_bs.size() != gadget->Dimension())
      return InvalidArgumentError(
          "`key_as` and `key_bs` must have the same length as the gadget dimension.");
    RnsGaloisKey key;
    RLWE_ASSIGN_OR_RETURN(key.d_as_, internal::PackKeys(key_as));
    RLWE_ASSIGN_OR_RETURN(key.d_bs_, internal::PackKeys(key_bs));
    RLWE_ASSIGN_OR_RETURN(key.ks_, internal::PrepareKey<ModularInt>(
                                       *key.d_as_, *key.d_bs_, (int)key_as.size(),
                                       ModuliSpan<ModularInt>(moduli.data(), moduli.size())));
    key.key_as_ = std::move(key_as);
    key.key_bs_ = std::move(key_bs);
    key.gadget_ = gadget;
    key.power_ = power;
    key.moduli_ = std::move(moduli);
    key.prng_seed_ = prng_seed;
    key.prng_type_ = prng_type;
    return key;
  }

  // rns/rns_galois_key.cc:204-236: key_b from the wire (unpacked on the
  // device), key_a regenerated from the seed.
  static StatusOr<RnsGaloisKey> Deserialize(const SerializedRnsGaloisKey& serialized,
                                            const RnsGadget<ModularInt>* gadget, ModuliVec moduli) {
    if (gadget == nullptr) return InvalidArgumentError("`gadget` must not be null.");
    const int dimension = serialized.key_bs_size();
    if (dimension <= 0) return InvalidArgumentError("`key_bs` must not be empty.");
    ModuliSpan<ModularInt> span(moduli.data(), moduli.size());
    std::vector<RnsPolynomial<ModularInt>> key_bs;
    for (int i = 0; i < dimension; ++i) {
      RLWE_ASSIGN_OR_RETURN(auto b, RnsPolynomial<ModularInt>::Deserialize(serialized.key_bs(i), span));
      key_bs.push_back(std::move(b));
    }
    RLWE_ASSIGN_OR_RETURN(auto key_as, SampleRandomPad(dimension, key_bs[0].LogN(), span,
                                                       serialized.prng_seed(),
                                                       serialized.prng_type()));
    return CreateFromKeyComponents(std::move(key_as), std::move(key_bs), serialized.power(),
                                   gadget, std::move(moduli), serialized.prng_seed(),
                                   serialized.prng_type());
  }

  // rns/rns_galois_key.h:129-137
  StatusOr<SerializedRnsGaloisKey> Serialize() const {
    SerializedRnsGaloisKey output;
    for (const auto& b : key_bs_) {
      RLWE_ASSIGN_OR_RETURN(*output.add_key_bs(), b.Serialize(Moduli()));
    }
    output.set_power(power_);
    output.set_prng_seed(prng_seed_);
    output.set_prng_type(prng_type_);
    return output;
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
    RLWE_ASSIGN_OR_RETURN(auto outs, SwitchPacked(*comps, (int)cts.size(), (mask >> 1) & 1,
                                                  cts[0]->LogN()));
    std::vector<RnsBgvCiphertext<ModularInt>> out;
    for (size_t c = 0; c < cts.size(); ++c) {
      std::vector<RnsPolynomial<ModularInt>> comps_new;
      comps_new.push_back(std::move(outs[2 * c]));
      comps_new.push_back(std::move(outs[2 * c + 1]));
      double error = cts[c]->Error() + cts[c]->ErrorParams()->BoundOnGadgetBasedKeySwitching(
                                           1, gadget_->LogGadgetBase(), gadget_->Dimension());
      out.emplace_back(std::move(comps_new), moduli_, 1, error, cts[c]->ErrorParams());
    }
    return out;
  }

  // rns/rns_galois_key.cc:289-308: sum_i digits[i] (.) a_i.
  StatusOr<RnsPolynomial<ModularInt>> PrecomputeRandomPad(
      const std::vector<RnsPolynomial<ModularInt>>& ciphertext_pad_digits) const {
    if ((int)ciphertext_pad_digits.size() != gadget_->Dimension())
      return InvalidArgumentError("`ciphertext_pad_digits` has incorrect size");
    return internal::LazyInnerProduct(ciphertext_pad_digits, *d_as_, Moduli());
  }

  // rns/rns_galois_key.h:181-193 with rns/rns_galois_key.cc:310-347: the
  // switched c0 from precomputed pad digits, c1 = the new pad.
  StatusOr<RnsBgvCiphertext<ModularInt>> ApplyToWithRandomPad(
      const RnsBgvCiphertext<ModularInt>& ciphertext,
      const std::vector<RnsPolynomial<ModularInt>>& ciphertext_pad_digits,
      RnsPolynomial<ModularInt> ciphertext_pad_new) const {
    if (ciphertext.PowerOfS() != power_)
      return InvalidArgumentError("`ciphertext` does not have a matching substitution power.");
    if (ciphertext.NumModuli() != (int)moduli_.size())
      return InvalidArgumentError("`ciphertext` does not have a matching RNS moduli set.");
    if (ciphertext.Degree() > 1)
      return InvalidArgumentError("Galois key can only apply to a ciphertext of degree 1.");
    if ((int)ciphertext_pad_digits.size() != gadget_->Dimension())
      return InvalidArgumentError("`ciphertext_pad_digits` has incorrect size");
    RLWE_ASSIGN_OR_RETURN(auto c0_new,
                          internal::LazyInnerProduct(ciphertext_pad_digits, *d_bs_, Moduli()));
    RLWE_ASSIGN_OR_RETURN(auto c0, ciphertext.Component(0));
    RLWE_RETURN_IF_ERROR(c0_new.AddInPlace(c0, Moduli()));
    double error = ciphertext.Error() + ciphertext.ErrorParams()->BoundOnGadgetBasedKeySwitching(
                                            1, gadget_->LogGadgetBase(), gadget_->Dimension());
    std::vector<RnsPolynomial<ModularInt>> comps;
    comps.push_back(std::move(c0_new));
    comps.push_back(std::move(ciphertext_pad_new));
    return RnsBgvCiphertext<ModularInt>(std::move(comps), moduli_, 1, error,
                                        ciphertext.ErrorParams());
  }

  // rns/rns_galois_key.cc:349-406: a key for s(X^(power * other.power)).
  // Column i of the composition is this key applied to the "ciphertext"
  // (sigma(b'_i), sigma(a'_i)), sigma = X -> X^power: decompose
  // INTT(sigma(a'_i)), multiply the transformed digits into a_j and b_j and
  // add sigma(b'_i) to the b side -- one batched Galois key switch of k
  // ciphertexts.
  StatusOr<RnsGaloisKey> Compose(const RnsGaloisKey& other_key) const {
    const int k = (int)other_key.key_as_.size();
    if (k != gadget_->Dimension())
      return InvalidArgumentError("`other_key` must have the gadget dimension of this key.");
    const int L = (int)moduli_.size();
    const int log_n = key_as_[0].LogN();
    const size_t words = (size_t)L << log_n;
    // (sigma(b'_i), sigma(a'_i)) back to back, as k degree-1 ciphertexts.
    device::Buffer cts(2 * words * (size_t)k);
    auto ids = DevicePrimes<ModularInt>(Moduli());
    for (int i = 0; i < k; ++i) {
      RLWE_RETURN_IF_ERROR(device::FromCode(shg_substitute(
          device::Context(), other_key.key_bs_[i].DeviceData(), cts.data() + 2 * i * words, power_,
          1, L, ids.data())));
      RLWE_RETURN_IF_ERROR(device::FromCode(shg_substitute(
          device::Context(), other_key.key_as_[i].DeviceData(),
          cts.data() + (2 * i + 1) * words, power_, 1, L, ids.data())));
    }
    RLWE_ASSIGN_OR_RETURN(auto outs, SwitchPacked(cts, k, 1, log_n));
    std::vector<RnsPolynomial<ModularInt>> composed_as, composed_bs;
    for (int i = 0; i < k; ++i) {
      composed_bs.push_back(std::move(outs[2 * i]));
      composed_as.push_back(std::move(outs[2 * i + 1]));
    }
    const int composed_power = (power_ * other_key.power_) % (1 << (log_n + 1));
    return CreateFromKeyComponents(std::move(composed_as), std::move(composed_bs), composed_power,
                                   other_key.gadget_, moduli_, std::string(), PRNG_TYPE_INVALID);
  }

  // rns/rns_galois_key.cc:408-430: sum_j gk1_as_digits[i][j] (.) gk0_as[j].
  static StatusOr<std::vector<RnsPolynomial<ModularInt>>> CreatePadOfComposedKey(
      const std::vector<RnsPolynomial<ModularInt>>& gk0_as,
      const std::vector<std::vector<RnsPolynomial<ModularInt>>>& gk1_as_digits,
      ModuliSpan<ModularInt> moduli) {
    RLWE_ASSIGN_OR_RETURN(auto packed, internal::PackKeys(gk0_as));
    std::vector<RnsPolynomial<ModularInt>> out;
    for (size_t i = 0; i < gk0_as.size(); ++i) {
      RLWE_ASSIGN_OR_RETURN(auto a, internal::LazyInnerProduct(gk1_as_digits[i], *packed, moduli));
      out.push_back(std::move(a));
    }
    return out;
  }

  // rns/rns_galois_key.cc:432-467: b_i = sigma(b'_i) + sum_j digits[i][j] (.) b_j.
  StatusOr<RnsGaloisKey> ComposeWithPads(
      const RnsGaloisKey& other_key,
      const std::vector<std::vector<RnsPolynomial<ModularInt>>>& other_a_digits,
      std::vector<RnsPolynomial<ModularInt>> composed_key_as) const {
    const int k = (int)other_key.key_bs_.size();
    std::vector<RnsPolynomial<ModularInt>> composed_bs;
    for (int i = 0; i < k; ++i) {
      RLWE_ASSIGN_OR_RETURN(auto b_sub, other_key.key_bs_[i].Substitute(power_, Moduli()));
      RLWE_ASSIGN_OR_RETURN(auto sum,
                            internal::LazyInnerProduct(other_a_digits[i], *d_bs_, Moduli()));
      RLWE_RETURN_IF_ERROR(b_sub.AddInPlace(sum, Moduli()));
      composed_bs.push_back(std::move(b_sub));
    }
    const int composed_power = (power_ * other_key.power_) % (1 << (composed_bs[0].LogN() + 1));
    return CreateFromKeyComponents(std::move(composed_key_as), std::move(composed_bs),
                                   composed_power, other_key.gadget_, moduli_, std::string(),
                                   PRNG_TYPE_INVALID);
  }

  const std::vector<RnsPolynomial<ModularInt>>& GetKeyA() const { return key_as_; }
  const std::vector<RnsPolynomial<ModularInt>>& GetKeyB() const { return key_bs_; }
  int SubstitutionPower() const { return power_; }
  int Dimension() const { return (int)key_as_.size(); }
  const RnsGadget<ModularInt>* Gadget() const { return gadget_; }
  const std::string& PrngSeed() const { return prng_seed_; }
  PrngType GetPrngType() const { return prng_type_; }

 private:
  RnsGaloisKey() = default;
  ModuliSpan<ModularInt> Moduli() const {
    return ModuliSpan<ModularInt>(moduli_.data(), moduli_.size());
  }
  // The fused Galois key switch of `count` packed degree-1 ciphertexts
  // (c0, c1 per ciphertext): returns out0, out1 of each, interleaved.
  StatusOr<std::vector<RnsPolynomial<ModularInt>>> SwitchPacked(const device::Buffer& comps,
                                                                int count, int c1_is_ntt,
                                                                int log_n) const {
    const int L = (int)moduli_.size();
    const size_t words = (size_t)L << log_n;
    device::Buffer o0(words * count), o1(words * count);
    auto ids = DevicePrimes<ModularInt>(Moduli());
    std::vector<int> lbs(gadget_->LogBases().begin(), gadget_->LogBases().end());
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_galois_apply_key(
        device::Context(), comps.data(), c1_is_ntt, ks_.get(), lbs.data(), o0.data(), o1.data(),
        count, L, ids.data())));
    std::vector<RnsPolynomial<ModularInt>> out;
    for (int c = 0; c < count; ++c) {
      out.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, o0.data() + c * words));
      out.push_back(RnsPolynomial<ModularInt>::FromDevice(log_n, L, true, o1.data() + c * words));
    }
    return out;
  }

  std::vector<RnsPolynomial<ModularInt>> key_as_, key_bs_;
  std::shared_ptr<device::Buffer> d_as_, d_bs_;
  std::shared_ptr<shg_ks_key> ks_;  // over d_as_, d_bs_ (declared after them)
  const RnsGadget<ModularInt>* gadget_ = nullptr;
  int power_ = 1;
  std::vector<const PrimeModulus<ModularInt>*> moduli_;
  std::string prng_seed_;
  PrngType prng_type_ = PRNG_TYPE_INVALID;
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_RELINEARIZATION_KEY_H_
