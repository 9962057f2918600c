// RnsPolynomial<MontgomeryInt<Uint64>> (rns/rns_polynomial.h:47-636) with
// its coefficient vectors resident in device memory.  Every arithmetic
// method runs on the GPU through libshellgpu; Coeffs() downloads a host
// mirror on demand.  Signatures, checks and error messages follow the
// reference; on error the polynomial is left unchanged.
#ifndef SHG_RLWE_RNS_POLYNOMIAL_H_
#define SHG_RLWE_RNS_POLYNOMIAL_H_

#include <memory>
#include <string>
#include <vector>

#include "rlwe/device.h"
#include "rlwe/prng.h"
#include "rlwe/rns_modulus.h"
#include "rlwe/serialization.h"

namespace rlwe {

template <typename ModularInt>
class RnsPolynomial {
 public:
  using ModularIntParams = typename ModularInt::Params;
  using NttParams = NttParameters<ModularInt>;
  using Moduli = ModuliSpan<ModularInt>;

  // rns/rns_polynomial.cc:34-52
  static StatusOr<RnsPolynomial> Create(std::vector<std::vector<ModularInt>> coeff_vectors,
                                        bool is_ntt = true) {
    if (coeff_vectors.empty())
      return InvalidArgumentError("`coeff_vectors` cannot be empty.");
    const size_t num_coeffs = coeff_vectors[0].size();
    if (num_coeffs == 0 || (num_coeffs & (num_coeffs - 1)) != 0)
      return InvalidArgumentError("`coeff_vectors` must contain vectors of length a power of two.");
    for (const auto& v : coeff_vectors)
      if (v.size() != num_coeffs)
        return InvalidArgumentError("`coeff_vectors` must contain vectors of equal length.");
    int log_n = 0;
    while ((size_t{1} << log_n) < num_coeffs) ++log_n;
    RnsPolynomial p(log_n, static_cast<int>(coeff_vectors.size()), is_ntt);
    std::vector<uint64_t> flat(p.Words());
    for (size_t i = 0; i < coeff_vectors.size(); ++i)
      for (size_t j = 0; j < num_coeffs; ++j)
        flat[i * num_coeffs + j] = coeff_vectors[i][j].GetMontgomeryRepresentation();
    RLWE_RETURN_IF_ERROR(p.buf_->Upload(flat.data(), flat.size()));
    return p;
  }

  static StatusOr<RnsPolynomial> CreateZero(int log_n, Moduli moduli, bool is_ntt = true) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
    RnsPolynomial p(log_n, static_cast<int>(moduli.size()), is_ntt);
    std::vector<uint64_t> flat(p.Words(), 0);
    RLWE_RETURN_IF_ERROR(p.buf_->Upload(flat.data(), flat.size()));
    return p;
  }

  static StatusOr<RnsPolynomial> CreateOne(int log_n, Moduli moduli) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
    RnsPolynomial p(log_n, static_cast<int>(moduli.size()), true);
    const size_t n = size_t{1} << log_n;
    std::vector<uint64_t> flat(p.Words());
    for (size_t i = 0; i < moduli.size(); ++i)
      for (size_t j = 0; j < n; ++j) flat[i * n + j] = moduli[i]->ModParams()->r_mod_modulus;
    RLWE_RETURN_IF_ERROR(p.buf_->Upload(flat.data(), flat.size()));
    return p;
  }

  // rns/rns_polynomial.h:640-667: ImportRandom (montgomery.h:260-273) for
  // every coefficient, limb by limb, from any SecurePrng: log2(q) random bits
  // per draw (a 64-bit draw, or one byte for moduli of 8 bits or fewer),
  // redrawn while >= q, stored as the representation.  The draws are made on
  // the host; RnsGaloisKey's seeded pads use shg_sample_uniform_chacha.
  template <typename Prng = SecurePrng>
  static StatusOr<RnsPolynomial> SampleUniform(int log_n, Prng* prng, Moduli moduli) {
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    if (prng == nullptr) return InvalidArgumentError("`prng` must not be null.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
    const size_t n = size_t{1} << log_n;
    RnsPolynomial p(log_n, static_cast<int>(moduli.size()), true);
    std::vector<uint64_t> flat(p.Words());
    for (size_t i = 0; i < moduli.size(); ++i) {
      const auto* mp = moduli[i]->ModParams();
      const int bits = (int)mp->log_modulus;
      for (size_t j = 0; j < n; ++j) {
        uint64_t r;
        do {
          r = 0;
          for (int need = bits; need > 0;) {
            if (need <= 8) {
              RLWE_ASSIGN_OR_RETURN(uint8_t v, prng->Rand8());
              r = (r << need) + (v & ((1u << need) - 1));
              break;
            }
            RLWE_ASSIGN_OR_RETURN(uint64_t v, prng->Rand64());
            const int take = need < 64 ? need : 64;
            r = (take == 64 ? 0 : (r << take)) + (take == 64 ? v : (v & ((1ull << take) - 1)));
            need -= take;
          }
        } while (r >= mp->modulus);
        flat[i * n + j] = r;
      }
    }
    RLWE_RETURN_IF_ERROR(p.buf_->Upload(flat.data(), flat.size()));
    return p;
  }

  // rns/rns_polynomial.cc:140-169: coefficients mod q in balanced form,
  // lifted to every modulus (ConvertModulusBalanced) and transformed.
  static StatusOr<RnsPolynomial> ConvertBalancedFromPolynomialCoeffs(
      const std::vector<ModularInt>& coeffs_q, const ModularIntParams* mod_params_q,
      Moduli moduli) {
    const size_t num_coeffs = coeffs_q.size();
    if (num_coeffs == 0 || (num_coeffs & (num_coeffs - 1)) != 0)
      return InvalidArgumentError("`coeffs_q` must have length a power of 2.");
    if (mod_params_q == nullptr) return InvalidArgumentError("`mod_params_q` must not be null.");
    if (moduli.empty()) return InvalidArgumentError("`moduli` must not be empty.");
    int log_n = 0;
    while ((size_t{1} << log_n) < num_coeffs) ++log_n;
    RnsPolynomial p(log_n, static_cast<int>(moduli.size()), false);
    std::vector<uint64_t> flat(p.Words());
    const uint64_t q = mod_params_q->modulus;
    for (size_t i = 0; i < moduli.size(); ++i) {
      const auto* mp = moduli[i]->ModParams();
      const uint64_t pi = mp->modulus;
      const uint64_t diff = q > pi ? q - pi : pi - q;
      RLWE_ASSIGN_OR_RETURN(ModularInt diff_p, ModularInt::ImportInt(diff, mp));
      for (size_t j = 0; j < num_coeffs; ++j) {
        if (pi == q) {
          flat[i * num_coeffs + j] = coeffs_q[j].GetMontgomeryRepresentation();
          continue;
        }
        const uint64_t x = coeffs_q[j].ExportInt(mod_params_q);
        RLWE_ASSIGN_OR_RETURN(ModularInt xp, ModularInt::ImportInt(x, mp));
        if (x > (q >> 1)) {
          if (pi > q) xp.AddInPlace(diff_p, mp); else xp.SubInPlace(diff_p, mp);
        }
        flat[i * num_coeffs + j] = xp.GetMontgomeryRepresentation();
      }
    }
    RLWE_RETURN_IF_ERROR(p.buf_->Upload(flat.data(), flat.size()));
    RLWE_RETURN_IF_ERROR(p.ConvertToNttForm(moduli));
    return p;
  }

  // rns/rns_polynomial.h:99-122: per limb MontgomeryInt::DeserializeVector
  // (montgomery.cc:184-219), unpacked on the device.
  static StatusOr<RnsPolynomial> Deserialize(const SerializedRnsPolynomial& serialized,
                                             Moduli moduli) {
    const int log_n = serialized.log_n();
    if (log_n <= 0) return InvalidArgumentError("`log_n` must be positive.");
    const int L = serialized.coeff_vectors_size();
    if (L != (int)moduli.size())
      return InvalidArgumentError("Number of serialized coefficient vectors must be " +
                                  std::to_string(moduli.size()));
    if (log_n > 15)
      return InvalidArgumentError("Number of coefficients, " + std::to_string(1LL << std::min(log_n, 62)) +
                                  ", cannot be larger than 32768.");
    RLWE_ASSIGN_OR_RETURN(std::vector<int> ids, DevicePrimesAt(moduli, log_n));
    std::vector<size_t> limb_bytes(L);
    size_t poly_bytes = 0;
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_serialized_bytes(
        device::Context(), ids.data(), L, limb_bytes.data(), &poly_bytes)));
    std::vector<uint64_t> staged((poly_bytes + 7) / 8 + 1, 0);
    char* dst = reinterpret_cast<char*>(staged.data());
    for (int i = 0; i < L; ++i) {
      const std::string& v = serialized.coeff_vectors(i);
      const long long bits = (long long)moduli[i]->ModParams()->log_modulus << log_n;
      if (v.size() < limb_bytes[i])
        return InvalidArgumentError("The input vector of size " + std::to_string(v.size()) +
                                    " is too small to contain " + std::to_string(bits) + " bits.");
      std::memcpy(dst, v.data(), limb_bytes[i]);
      dst += limb_bytes[i];
    }
    device::Buffer bytes(staged.size());
    RLWE_RETURN_IF_ERROR(bytes.Upload(staged.data(), staged.size()));
    RnsPolynomial p(log_n, L, serialized.is_ntt());
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_deserialize_polys(
        device::Context(), reinterpret_cast<const uint8_t*>(bytes.data()), 1, L, ids.data(),
        p.buf_->data())));
    return p;
  }

  // rns/rns_polynomial.h:124-140: per limb MontgomeryInt::SerializeVector,
  // packed on the device.
  StatusOr<SerializedRnsPolynomial> Serialize(Moduli moduli) const {
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    RLWE_ASSIGN_OR_RETURN(std::vector<int> ids, DevicePrimesAt(moduli, log_n_));
    std::vector<size_t> limb_bytes(num_moduli_);
    size_t poly_bytes = 0;
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_serialized_bytes(
        device::Context(), ids.data(), num_moduli_, limb_bytes.data(), &poly_bytes)));
    device::Buffer bytes((poly_bytes + 7) / 8);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_serialize_polys(
        device::Context(), buf_->data(), 1, num_moduli_, ids.data(),
        reinterpret_cast<uint8_t*>(bytes.data()))));
    std::vector<uint64_t> host(bytes.words());
    RLWE_RETURN_IF_ERROR(bytes.Download(host.data(), host.size()));
    const char* src = reinterpret_cast<const char*>(host.data());
    SerializedRnsPolynomial out;
    out.set_log_n(log_n_);
    out.set_is_ntt(is_ntt_);
    for (int i = 0; i < num_moduli_; ++i) {
      out.add_coeff_vectors()->assign(src, limb_bytes[i]);
      src += limb_bytes[i];
    }
    return out;
  }

  // Wraps an existing device buffer region (used by batched containers).
  static RnsPolynomial FromDevice(int log_n, int num_moduli, bool is_ntt, const uint64_t* src) {
    RnsPolynomial p(log_n, num_moduli, is_ntt);
    (void)device::CopyWords(p.buf_->data(), src, p.Words());
    return p;
  }

  RnsPolynomial(const RnsPolynomial& that)
      : log_n_(that.log_n_), num_moduli_(that.num_moduli_), is_ntt_(that.is_ntt_),
        buf_(std::make_shared<device::Buffer>(that.Words())) {
    (void)device::CopyWords(buf_->data(), that.buf_->data(), Words());
  }
  RnsPolynomial& operator=(const RnsPolynomial& that) {
    if (this == &that) return *this;
    RnsPolynomial tmp(that);
    *this = std::move(tmp);
    return *this;
  }
  RnsPolynomial(RnsPolynomial&&) = default;
  RnsPolynomial& operator=(RnsPolynomial&&) = default;

  // rns/rns_polynomial.cc:171-188
  Status ConvertToNttForm(Moduli moduli) {
    if (is_ntt_) return InvalidArgumentError("Polynomial already in NTT form");
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(
        shg_ntt_fwd(device::Context(), buf_->data(), 1, num_moduli_, ids.data(), SHG_NTT_AUTO)));
    is_ntt_ = true;
    Touch();
    return OkStatus();
  }
  // rns/rns_polynomial.cc:190-207
  Status ConvertToCoeffForm(Moduli moduli) {
    if (!is_ntt_) return InvalidArgumentError("Polynomial already in Coefficient form");
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(
        shg_ntt_inv(device::Context(), buf_->data(), 1, num_moduli_, ids.data(), SHG_NTT_AUTO)));
    is_ntt_ = false;
    Touch();
    return OkStatus();
  }

  StatusOr<RnsPolynomial> Negate(Moduli moduli) const {
    RnsPolynomial out = *this;
    RLWE_RETURN_IF_ERROR(out.NegateInPlace(moduli));
    return out;
  }
  Status NegateInPlace(Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    return Pointwise(SHG_OP_NEG, nullptr, nullptr, moduli);
  }
  StatusOr<RnsPolynomial> Add(const RnsPolynomial& that, Moduli moduli) const {
    RnsPolynomial out = *this;
    RLWE_RETURN_IF_ERROR(out.AddInPlace(that, moduli));
    return out;
  }
  StatusOr<RnsPolynomial> Sub(const RnsPolynomial& that, Moduli moduli) const {
    RnsPolynomial out = *this;
    RLWE_RETURN_IF_ERROR(out.SubInPlace(that, moduli));
    return out;
  }
  Status AddInPlace(const RnsPolynomial& that, Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckThat(that, "that"));
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    return Pointwise(SHG_OP_ADD, &that, nullptr, moduli);
  }
  Status SubInPlace(const RnsPolynomial& that, Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckThat(that, "that"));
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    return Pointwise(SHG_OP_SUB, &that, nullptr, moduli);
  }
  StatusOr<RnsPolynomial> Mul(typename ModularInt::Int scalar, Moduli moduli) const {
    RnsPolynomial out = *this;
    RLWE_RETURN_IF_ERROR(out.MulInPlace(scalar, moduli));
    return out;
  }
  Status MulInPlace(typename ModularInt::Int scalar, Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(
        shg_mul_int(device::Context(), buf_->data(), scalar, 1, num_moduli_, ids.data())));
    Touch();
    return OkStatus();
  }
  Status MulInPlace(const RnsInt<ModularInt>& scalar, Moduli moduli) {
    return MulInPlace(Span<const ModularInt>(scalar.RnsRep().data(), scalar.RnsRep().size()),
                      moduli);
  }
  Status MulInPlace(Span<const ModularInt> scalar_mod_qs, Moduli moduli) {
    if ((int)scalar_mod_qs.size() != num_moduli_)
      return InvalidArgumentError("`scalar_mod_qs` must contain " + std::to_string(num_moduli_) +
                                  " modular integers.");
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    std::vector<uint64_t> s(num_moduli_);
    for (int i = 0; i < num_moduli_; ++i) s[i] = scalar_mod_qs[i].GetMontgomeryRepresentation();
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_mul_scalar(device::Context(), buf_->data(), s.data(),
                                                         1, num_moduli_, ids.data())));
    Touch();
    return OkStatus();
  }
  StatusOr<RnsPolynomial> Mul(const RnsPolynomial& that, Moduli moduli) const {
    RnsPolynomial out = *this;
    RLWE_RETURN_IF_ERROR(out.MulInPlace(that, moduli));
    return out;
  }
  Status MulInPlace(const RnsPolynomial& that, Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckThat(that, "that"));
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    if (!is_ntt_) return InvalidArgumentError("RNS polynomial `this` must be in NTT form.");
    if (!that.is_ntt_) return InvalidArgumentError("RNS polynomial `that` must be in NTT form.");
    return Pointwise(SHG_OP_MUL, &that, nullptr, moduli);
  }
  Status FusedMulAddInPlace(const RnsPolynomial& a, const RnsPolynomial& b, Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckThat(a, "a"));
    RLWE_RETURN_IF_ERROR(CheckThat(b, "b"));
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    if (!is_ntt_) return InvalidArgumentError("RNS polynomial `this` must be in NTT form.");
    if (!a.is_ntt_) return InvalidArgumentError("RNS polynomial `a` must be in NTT form.");
    if (!b.is_ntt_) return InvalidArgumentError("RNS polynomial `b` must be in NTT form.");
    return Pointwise(SHG_OP_FMA, &a, &b, moduli);
  }

  // rns/rns_polynomial.cc:209-260
  StatusOr<RnsPolynomial> Substitute(int power, Moduli moduli) const {
    const int n = NumCoeffs();
    if (power < 0 || (power % 2) == 0 || power >= 2 * n)
      return InvalidArgumentError(
          "Substitution power must be a non-negative odd integer less than 2*n.");
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    if (!is_ntt_) return InvalidArgumentError("RnsPolynomial must be in NTT form");
    RnsPolynomial out(log_n_, num_moduli_, is_ntt_);
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_substitute(
        device::Context(), buf_->data(), out.buf_->data(), power, 1, num_moduli_, ids.data())));
    return out;
  }

  // rns/rns_polynomial.cc:262-344
  Status ModReduceLsb(typename ModularInt::Int t, const RnsInt<ModularInt>& ql_inv,
                      Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    if (num_moduli_ <= 1)
      return InvalidArgumentError("ModReduceLsb cannot apply with only one prime modulus.");
    if (ql_inv.NumModuli() != num_moduli_ - 1)
      return InvalidArgumentError("`ql_inv` must be defined with respect to " +
                                  std::to_string(num_moduli_ - 1) + " moduli.");
    std::vector<uint64_t> qi(num_moduli_ - 1);
    for (int i = 0; i < num_moduli_ - 1; ++i)
      qi[i] = ql_inv.zs[i].GetMontgomeryRepresentation();
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_mod_reduce_lsb(
        device::Context(), buf_->data(), t, qi.data(), is_ntt_ ? 1 : 0, 1, num_moduli_,
        ids.data())));
    // The result occupies the first num_moduli_-1 limbs of the buffer.
    num_moduli_ -= 1;
    Touch();
    return OkStatus();
  }

  // rns/rns_polynomial.cc:349-417: (a - lift([a]_{q_l})) / q_l, the BFV
  // modulus switch, in either form.
  Status ModReduceMsb(const RnsInt<ModularInt>& ql_inv, Moduli moduli) {
    RLWE_RETURN_IF_ERROR(CheckModuli(moduli));
    if (num_moduli_ <= 1)
      return InvalidArgumentError("ModReduceMsb cannot apply with only one prime modulus.");
    if (ql_inv.NumModuli() != num_moduli_ - 1)
      return InvalidArgumentError("`ql_inv` must be defined with respect to " +
                                  std::to_string(num_moduli_ - 1) + " moduli.");
    std::vector<uint64_t> qi(num_moduli_ - 1);
    for (int i = 0; i < num_moduli_ - 1; ++i) qi[i] = ql_inv.zs[i].GetMontgomeryRepresentation();
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_mod_reduce_msb(
        device::Context(), buf_->data(), qi.data(), is_ntt_ ? 1 : 0, 1, num_moduli_, ids.data())));
    num_moduli_ -= 1;
    Touch();
    return OkStatus();
  }

  // rns/rns_polynomial.cc:419-517: exact conversion to `output_moduli`
  // (the centered representative), coefficient form.
  StatusOr<RnsPolynomial> SwitchRnsBasis(Moduli this_moduli, Moduli output_moduli,
                                         Span<const ModularInt> prime_q_hat_inv_mod_qs,
                                         Span<const RnsInt<ModularInt>> prime_q_hat_mod_ps,
                                         Span<const ModularInt> q_mod_ps) const {
    if (is_ntt_) return InvalidArgumentError("Polynomial must be in coefficient form.");
    RLWE_RETURN_IF_ERROR(CheckBasisArgs(this_moduli, output_moduli, prime_q_hat_inv_mod_qs.size(),
                                        prime_q_hat_mod_ps.size(), false));
    if (q_mod_ps.size() != output_moduli.size())
      return InvalidArgumentError("`q_mod_ps` must contain " + std::to_string(output_moduli.size()) +
                                  " elements.");
    std::vector<uint64_t> mat, vec;
    for (int i = 0; i < num_moduli_; ++i)
      for (size_t j = 0; j < output_moduli.size(); ++j)
        mat.push_back(prime_q_hat_mod_ps[i].Component(j).GetMontgomeryRepresentation());
    for (size_t j = 0; j < output_moduli.size(); ++j) vec.push_back(q_mod_ps[j].GetMontgomeryRepresentation());
    return Switch(SHG_BASIS_EXACT, false, this_moduli, output_moduli, prime_q_hat_inv_mod_qs, mat,
                  vec);
  }

  // rns/rns_polynomial.cc:519-636: round(P * a / Q) mod P, coefficient form.
  StatusOr<RnsPolynomial> ScaleAndSwitchRnsBasis(Moduli this_moduli, Moduli output_moduli,
                                                 Span<const ModularInt> prime_q_hat_inv_mod_qs,
                                                 Span<const RnsInt<ModularInt>> prime_q_inv_mod_ps,
                                                 Span<const ModularInt> p_mod_qs) const {
    if (is_ntt_) return InvalidArgumentError("Polynomial must be in coefficient form.");
    RLWE_RETURN_IF_ERROR(CheckBasisArgs(this_moduli, output_moduli, prime_q_hat_inv_mod_qs.size(),
                                        num_moduli_, false));
    if (prime_q_inv_mod_ps.size() != output_moduli.size())
      return InvalidArgumentError("`prime_q_inv_mod_ps` must contain " +
                                  std::to_string(output_moduli.size()) + " elements.");
    if ((int)p_mod_qs.size() < num_moduli_)
      return InvalidArgumentError("`p_mod_qs` must contain at least " +
                                  std::to_string(num_moduli_) + " elements.");
    std::vector<uint64_t> mat, vec;
    for (size_t j = 0; j < output_moduli.size(); ++j)
      for (int i = 0; i < num_moduli_; ++i)
        mat.push_back(prime_q_inv_mod_ps[j].Component(i).GetMontgomeryRepresentation());
    for (int i = 0; i < num_moduli_; ++i) vec.push_back(p_mod_qs[i].GetMontgomeryRepresentation());
    return Switch(SHG_BASIS_SCALE, false, this_moduli, output_moduli, prime_q_hat_inv_mod_qs, mat,
                  vec);
  }

  // rns/rns_polynomial.cc:638-701: fast approximate conversion, input in
  // either form, coefficient-form output.
  StatusOr<RnsPolynomial> ApproxSwitchRnsBasis(Moduli this_moduli, Moduli output_moduli,
                                               Span<const ModularInt> prime_q_hat_inv_mod_qs,
                                               Span<const RnsInt<ModularInt>> prime_q_hat_mod_ps,
                                               bool is_balanced_rep = false) const {
    RLWE_RETURN_IF_ERROR(CheckBasisArgs(this_moduli, output_moduli, prime_q_hat_inv_mod_qs.size(),
                                        prime_q_hat_mod_ps.size(), true));
    std::vector<uint64_t> mat;
    for (int i = 0; i < num_moduli_; ++i)
      for (size_t j = 0; j < output_moduli.size(); ++j)
        mat.push_back(prime_q_hat_mod_ps[i].Component(j).GetMontgomeryRepresentation());
    return Switch(SHG_BASIS_APPROX, is_balanced_rep, this_moduli, output_moduli,
                  prime_q_hat_inv_mod_qs, mat, {});
  }

  // rns/rns_polynomial.cc:703-797: P^-1 (p_mod_t * a + t [k * aux]_P) mod Q
  // with k = -p_mod_t * t^-1 mod P, delta switched to Q approximately with
  // the balanced lift.
  Status ApproxModReduceLsb(RnsPolynomial polynomial_aux, Moduli this_moduli, Moduli aux_moduli,
                            Span<const ModularInt> p_inv_mod_qs,
                            Span<const ModularInt> prime_p_hat_inv_mod_ps,
                            Span<const RnsInt<ModularInt>> prime_p_hat_mod_qs,
                            Span<const ModularInt> t_inv_mod_ps, typename ModularInt::Int p_mod_t,
                            typename ModularInt::Int t) {
    RLWE_RETURN_IF_ERROR(CheckApproxModReduce(polynomial_aux, this_moduli, aux_moduli, p_inv_mod_qs,
                                              prime_p_hat_inv_mod_ps, prime_p_hat_mod_qs));
    if ((int)t_inv_mod_ps.size() != polynomial_aux.num_moduli_)
      return InvalidArgumentError("`t_inv_mod_ps` must contain " +
                                  std::to_string(polynomial_aux.num_moduli_) + " elements.");
    // aux <- aux * p_mod_t * (-t^-1) per aux limb.
    std::vector<ModularInt> scal;
    for (size_t j = 0; j < aux_moduli.size(); ++j) {
      const auto* mp = aux_moduli[j]->ModParams();
      RLWE_ASSIGN_OR_RETURN(ModularInt r, ModularInt::ImportInt(p_mod_t, mp));
      ModularInt neg = ModularInt(mp->modulus - t_inv_mod_ps[j].GetMontgomeryRepresentation());
      scal.push_back(r.Mul(neg, mp));
    }
    RLWE_RETURN_IF_ERROR(polynomial_aux.MulInPlace(Span<const ModularInt>(scal), aux_moduli));
    RLWE_ASSIGN_OR_RETURN(RnsPolynomial delta,
                          polynomial_aux.ApproxSwitchRnsBasis(aux_moduli, this_moduli,
                                                              prime_p_hat_inv_mod_ps,
                                                              prime_p_hat_mod_qs, true));
    RLWE_RETURN_IF_ERROR(delta.ConvertToNttForm(this_moduli));
    RLWE_RETURN_IF_ERROR(delta.MulInPlace(t, this_moduli));
    if (!is_ntt_) RLWE_RETURN_IF_ERROR(ConvertToNttForm(this_moduli));
    RLWE_RETURN_IF_ERROR(MulInPlace(p_mod_t, this_moduli));
    RLWE_RETURN_IF_ERROR(AddInPlace(delta, this_moduli));
    return MulInPlace(p_inv_mod_qs, this_moduli);
  }

  // rns/rns_polynomial.cc:799-852: P^-1 (a - [aux]_P) mod Q.
  Status ApproxModReduceMsb(RnsPolynomial polynomial_aux, Moduli this_moduli, Moduli aux_moduli,
                            Span<const ModularInt> p_inv_mod_qs,
                            Span<const ModularInt> prime_p_hat_inv_mod_ps,
                            Span<const RnsInt<ModularInt>> prime_p_hat_mod_qs) {
    RLWE_RETURN_IF_ERROR(CheckApproxModReduce(polynomial_aux, this_moduli, aux_moduli, p_inv_mod_qs,
                                              prime_p_hat_inv_mod_ps, prime_p_hat_mod_qs));
    RLWE_ASSIGN_OR_RETURN(RnsPolynomial delta,
                          polynomial_aux.ApproxSwitchRnsBasis(aux_moduli, this_moduli,
                                                              prime_p_hat_inv_mod_ps,
                                                              prime_p_hat_mod_qs, false));
    RLWE_RETURN_IF_ERROR(delta.ConvertToNttForm(this_moduli));
    if (!is_ntt_) RLWE_RETURN_IF_ERROR(ConvertToNttForm(this_moduli));
    RLWE_RETURN_IF_ERROR(SubInPlace(delta, this_moduli));
    return MulInPlace(p_inv_mod_qs, this_moduli);
  }

  void AppendCoeffVector(std::vector<ModularInt> coeffs) {
    auto host = Coeffs();
    host.push_back(std::move(coeffs));
    *this = Create(std::move(host), is_ntt_).value();
  }
  StatusOr<std::vector<ModularInt>> DetachLastCoeffVector() {
    if (num_moduli_ == 0) return InvalidArgumentError("RnsPolynomial is empty.");
    std::vector<ModularInt> last = Coeffs().back();
    num_moduli_ -= 1;
    Touch();
    return last;
  }

  bool operator==(const RnsPolynomial& that) const {
    return log_n_ == that.log_n_ && is_ntt_ == that.is_ntt_ && Coeffs() == that.Coeffs();
  }
  bool operator!=(const RnsPolynomial& that) const { return !(*this == that); }

  int LogN() const { return log_n_; }
  int NumCoeffs() const { return 1 << log_n_; }
  int NumModuli() const { return num_moduli_; }
  bool IsNttForm() const { return is_ntt_; }
  RnsPolynomial Clone() const { return *this; }

  // Host mirror of the coefficient vectors (downloaded when stale).
  const std::vector<std::vector<ModularInt>>& Coeffs() const {
    if (!host_valid_) {
      const size_t n = NumCoeffs();
      std::vector<uint64_t> flat((size_t)num_moduli_ * n);
      if (!flat.empty()) (void)buf_->Download(flat.data(), flat.size());
      host_.clear();
      for (int i = 0; i < num_moduli_; ++i) {
        std::vector<ModularInt> v;
        v.reserve(n);
        for (size_t j = 0; j < n; ++j) v.emplace_back(flat[i * n + j]);
        host_.push_back(std::move(v));
      }
      host_valid_ = true;
    }
    return host_;
  }

  // Device access for batched containers and the key-switching classes.
  uint64_t* DeviceData() { Touch(); return buf_->data(); }
  const uint64_t* DeviceData() const { return buf_->data(); }

 private:
  RnsPolynomial(int log_n, int num_moduli, bool is_ntt)
      : log_n_(log_n), num_moduli_(num_moduli), is_ntt_(is_ntt),
        buf_(std::make_shared<device::Buffer>((size_t)num_moduli << log_n)) {}

  size_t Words() const { return (size_t)num_moduli_ << log_n_; }
  void Touch() { host_valid_ = false; }

  Status CheckModuli(Moduli moduli) const {
    if ((int)moduli.size() != num_moduli_)
      return InvalidArgumentError("`moduli` must contain " + std::to_string(num_moduli_) +
                                  " RNS moduli.");
    return OkStatus();
  }
  Status CheckBasisArgs(Moduli this_moduli, Moduli output_moduli, size_t num_q_hat_inv,
                        size_t num_q_hat, bool exact_counts) const {
    if ((int)this_moduli.size() != num_moduli_)
      return InvalidArgumentError("`this_moduli` must contain " + std::to_string(num_moduli_) +
                                  " RNS moduli.");
    if (output_moduli.empty()) return InvalidArgumentError("`output_moduli` must not be empty.");
    const std::string qualifier = exact_counts ? "" : "at least ";
    if (exact_counts ? num_q_hat_inv != (size_t)num_moduli_ : num_q_hat_inv < (size_t)num_moduli_)
      return InvalidArgumentError("`prime_q_hat_inv_mod_qs` must contain " + qualifier +
                                  std::to_string(num_moduli_) + " elements.");
    if (exact_counts ? num_q_hat != (size_t)num_moduli_ : num_q_hat < (size_t)num_moduli_)
      return InvalidArgumentError("`prime_q_hat_mod_ps` must contain " + qualifier +
                                  std::to_string(num_moduli_) + " elements.");
    return OkStatus();
  }
  Status CheckApproxModReduce(const RnsPolynomial& aux, Moduli this_moduli, Moduli aux_moduli,
                              Span<const ModularInt> p_inv_mod_qs,
                              Span<const ModularInt> prime_p_hat_inv_mod_ps,
                              Span<const RnsInt<ModularInt>> prime_p_hat_mod_qs) const {
    if ((int)this_moduli.size() != num_moduli_)
      return InvalidArgumentError("`this_moduli` must contain " + std::to_string(num_moduli_) +
                                  " RNS moduli.");
    if ((int)aux_moduli.size() != aux.num_moduli_)
      return InvalidArgumentError("`aux_moduli` must contain " + std::to_string(aux.num_moduli_) +
                                  " RNS moduli.");
    if ((int)p_inv_mod_qs.size() != num_moduli_)
      return InvalidArgumentError("`p_inv_mod_qs` must contain " + std::to_string(num_moduli_) +
                                  " elements.");
    if ((int)prime_p_hat_inv_mod_ps.size() != aux.num_moduli_)
      return InvalidArgumentError("`prime_p_hat_inv_mod_ps` must contain " +
                                  std::to_string(aux.num_moduli_) + " elements.");
    if ((int)prime_p_hat_mod_qs.size() != aux.num_moduli_)
      return InvalidArgumentError("`prime_p_hat_mod_qs` must contain " +
                                  std::to_string(aux.num_moduli_) + " elements.");
    return OkStatus();
  }
  StatusOr<RnsPolynomial> Switch(int mode, bool balanced, Moduli this_moduli, Moduli output_moduli,
                                 Span<const ModularInt> q_hat_inv, const std::vector<uint64_t>& mat,
                                 const std::vector<uint64_t>& vec) const {
    std::vector<uint64_t> qh;
    for (int i = 0; i < num_moduli_; ++i) qh.push_back(q_hat_inv[i].GetMontgomeryRepresentation());
    auto in_ids = DevicePrimes<ModularInt>(this_moduli);
    auto out_ids = DevicePrimes<ModularInt>(output_moduli);
    RnsPolynomial out(log_n_, (int)output_moduli.size(), false);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_switch_rns_basis(
        device::Context(), mode, buf_->data(), is_ntt_ ? 1 : 0, balanced ? 1 : 0, 1, num_moduli_,
        in_ids.data(), out.buf_->data(), (int)output_moduli.size(), out_ids.data(), qh.data(),
        mat.data(), vec.empty() ? nullptr : vec.data())));
    return out;
  }
  Status CheckThat(const RnsPolynomial& that, const char* name) const {
    if (that.num_moduli_ != num_moduli_)
      return InvalidArgumentError(std::string("`") + name + "` must contain " +
                                  std::to_string(num_moduli_) + " coefficient vectors.");
    return OkStatus();
  }
  Status Pointwise(int op, const RnsPolynomial* a, const RnsPolynomial* b, Moduli moduli) {
    auto ids = DevicePrimes<ModularInt>(moduli);
    RLWE_RETURN_IF_ERROR(device::FromCode(shg_pointwise(
        device::Context(), op, buf_->data(), a ? a->buf_->data() : nullptr, 0,
        b ? b->buf_->data() : nullptr, 0, 1, num_moduli_, ids.data())));
    Touch();
    return OkStatus();
  }

  int log_n_;
  int num_moduli_;
  bool is_ntt_;
  std::shared_ptr<device::Buffer> buf_;
  mutable std::vector<std::vector<ModularInt>> host_;
  mutable bool host_valid_ = false;
};

}  // namespace rlwe

#endif  // SHG_RLWE_RNS_POLYNOMIAL_H_
