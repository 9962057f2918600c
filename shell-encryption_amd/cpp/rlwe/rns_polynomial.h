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
#include "rlwe/rns_modulus.h"

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
