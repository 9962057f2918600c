// MontgomeryIntParams<Uint64> and MontgomeryInt<Uint64>: host-side scalar
// Montgomery arithmetic with the reference's constants, member names and
// semantics (shell_encryption/montgomery.h:77-666, montgomery.cc:26-499).
// The GPU kernels receive these constants; bulk arithmetic runs on the device.
#ifndef SHG_RLWE_MONTGOMERY_H_
#define SHG_RLWE_MONTGOMERY_H_

#include <cstdint>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "rlwe/status.h"

namespace rlwe {

using Uint64 = uint64_t;
using Uint128 = unsigned __int128;

template <typename T>
struct MontgomeryIntParams;

// Only the 64-bit instantiation is on the GPU path: residues are stored in
// one 64-bit word and products use 128-bit intermediates.
template <>
struct MontgomeryIntParams<Uint64> {
  using Int = Uint64;
  using BigInt = Uint128;
  static const size_t bitsize_int = 64;
  static const size_t bitsize_bigint = 128;

  // montgomery.cc:44-59: rejects q >= 2^62 and even q.
  static StatusOr<std::unique_ptr<const MontgomeryIntParams>> Create(Int modulus) {
    if ((modulus >> (bitsize_int - 2)) != 0) {
      return InvalidArgumentError("The modulus should be less than 2^62.");
    }
    if ((modulus % 2) == 0) {
      return InvalidArgumentError("The modulus should be odd.");
    }
    return std::unique_ptr<const MontgomeryIntParams>(new MontgomeryIntParams(modulus));
  }

  const BigInt r = static_cast<BigInt>(1) << bitsize_int;
  const Int modulus;
  const BigInt modulus_bigint;
  const unsigned int log_modulus;
  const Int r_mod_modulus;
  const Int r_mod_modulus_barrett;
  const Int inv_modulus;  // R*inv_r - modulus*inv_modulus = 1
  const Int inv_r;
  const Int inv_r_barrett;
  const BigInt barrett_numerator;         // 2^64 / modulus
  const BigInt barrett_numerator_bigint;  // 2^127 / modulus

  Int Zero() const { return 0; }
  Int One() const { return 1; }
  Int Two() const { return 2; }

  // montgomery.h:134-141
  Int BarrettReduce(Int input) const {
    Int out = static_cast<Int>((barrett_numerator * input) >> bitsize_int);
    out = input - out * modulus;
    return (out >= modulus) ? out - modulus : out;
  }

  // montgomery.h:143-152: reduction of a 128-bit value, using the 256-bit
  // product floor(barrett_numerator_bigint * input / 2^127).
  Int BarrettReduceBigInt(BigInt input) const {
    const Uint64 a0 = static_cast<Uint64>(barrett_numerator_bigint);
    const Uint64 a1 = static_cast<Uint64>(barrett_numerator_bigint >> 64);
    const Uint64 b0 = static_cast<Uint64>(input);
    const Uint64 b1 = static_cast<Uint64>(input >> 64);
    const Uint128 p00 = static_cast<Uint128>(a0) * b0;
    const Uint128 p01 = static_cast<Uint128>(a0) * b1;
    const Uint128 p10 = static_cast<Uint128>(a1) * b0;
    const Uint128 p11 = static_cast<Uint128>(a1) * b1;
    const Uint128 mid = (p00 >> 64) + static_cast<Uint64>(p01) + static_cast<Uint64>(p10);
    const Uint128 hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
    const Uint64 w1 = static_cast<Uint64>(mid);
    Int out = static_cast<Int>((hi << 1) | (w1 >> 63));
    out = static_cast<Int>(input) - out * modulus;
    return (out >= modulus) ? out - modulus : out;
  }

  // montgomery.h:156-164
  Int ExportInt(Int input) const {
    Int out = static_cast<Int>((static_cast<BigInt>(inv_r_barrett) * input) >> bitsize_int);
    out = input * inv_r - out * modulus;
    return (out >= modulus) ? out - modulus : out;
  }

  unsigned int SerializedSize() const { return (log_modulus + 7) / 8; }
  static bool DoesLogNFit(Uint64 log_n) { return log_n < bitsize_int - 1; }
  static double GetDouble(Int value) { return static_cast<double>(value); }

 private:
  static unsigned BitLength(Int x) {
    unsigned b = 0;
    while (x) { ++b; x >>= 1; }
    return b;
  }
  // montgomery.cc:62-92 (Hacker's Delight): x*2^w - y*modulus stays a power
  // of two while halving from 2^w down to 1.
  static std::tuple<Int, Int> Inverses(BigInt modulus_bigint, BigInt r) {
    BigInt x = 1, y = 0;
    for (int i = bitsize_int; i > 0; i--) {
      if ((x & 1) == 1) {
        x += modulus_bigint;
        y += r;
      }
      x >>= 1;
      y >>= 1;
    }
    return std::make_tuple(static_cast<Int>(x), static_cast<Int>(y));
  }
  explicit MontgomeryIntParams(Int mod)
      : modulus(mod),
        modulus_bigint(static_cast<BigInt>(mod)),
        log_modulus(BitLength(mod)),
        r_mod_modulus(static_cast<Int>(r % modulus_bigint)),
        r_mod_modulus_barrett(
            static_cast<Int>((static_cast<BigInt>(r_mod_modulus) << bitsize_int) / mod)),
        inv_modulus(std::get<1>(Inverses(modulus_bigint, r))),
        inv_r(std::get<0>(Inverses(modulus_bigint, r))),
        inv_r_barrett(static_cast<Int>((static_cast<BigInt>(inv_r) << bitsize_int) / mod)),
        barrett_numerator(r / modulus_bigint),
        barrett_numerator_bigint((static_cast<BigInt>(1) << 127) / modulus_bigint) {}
};

template <typename T>
class MontgomeryInt;

// A residue in Montgomery representation n = x * 2^64 mod q.
template <>
class MontgomeryInt<Uint64> {
 public:
  using Int = Uint64;
  using BigInt = Uint128;
  using Params = MontgomeryIntParams<Uint64>;

  explicit MontgomeryInt(Int n) : n_(n) {}
  MontgomeryInt(const MontgomeryInt&) = default;
  MontgomeryInt& operator=(const MontgomeryInt&) = default;

  // ToMontgomeryInt (montgomery.cc:26-40): Shoup product by R mod q.
  static StatusOr<MontgomeryInt> ImportInt(Int n, const Params* params) {
    BigInt product = static_cast<BigInt>(params->r_mod_modulus_barrett) * n;
    Int result = static_cast<Int>(product >> 64);
    result = n * params->r_mod_modulus - result * params->modulus;
    result -= (result >= params->modulus) ? params->modulus : 0;
    return MontgomeryInt(result);
  }
  static StatusOr<std::vector<MontgomeryInt>> BatchImportInts(const std::vector<Int>& ints,
                                                               const Params* params) {
    std::vector<MontgomeryInt> out;
    out.reserve(ints.size());
    for (Int v : ints) out.push_back(ImportInt(v, params).value());
    return out;
  }
  static MontgomeryInt ImportZero(const Params* params) { return MontgomeryInt(0); }
  static MontgomeryInt ImportOne(const Params* params) {
    return MontgomeryInt(params->r_mod_modulus);
  }

  Int ExportInt(const Params* params) const { return params->ExportInt(n_); }
  Int GetMontgomeryRepresentation() const { return n_; }

  // montgomery.h:318-336 (REDC)
  MontgomeryInt& MulInPlace(const MontgomeryInt& that, const Params* params) {
    BigInt product = static_cast<BigInt>(n_) * that.n_;
    Int u = static_cast<Int>(product) * params->inv_modulus;
    BigInt t = product + params->modulus_bigint * u;
    Int t_msb = static_cast<Int>(t >> 64);
    n_ = (t_msb >= params->modulus) ? t_msb - params->modulus : t_msb;
    return *this;
  }
  MontgomeryInt Mul(const MontgomeryInt& that, const Params* params) const {
    MontgomeryInt out(*this);
    return out.MulInPlace(that, params);
  }
  // montgomery.cc:221-227
  std::tuple<Int, Int> GetConstant(const Params* params) const {
    Int constant = ExportInt(params);
    Int constant_barrett =
        static_cast<Int>((static_cast<BigInt>(constant) << 64) / params->modulus);
    return std::make_tuple(constant, constant_barrett);
  }
  // montgomery.h:362-372 (Shoup)
  MontgomeryInt& MulConstantInPlace(const Int& constant, const Int& constant_barrett,
                                    const Params* params) {
    Int out = static_cast<Int>((static_cast<BigInt>(constant_barrett) * n_) >> 64);
    n_ = n_ * constant - out * params->modulus;
    n_ -= (n_ >= params->modulus) ? params->modulus : 0;
    return *this;
  }
  MontgomeryInt MulConstant(const Int& c, const Int& cb, const Params* params) const {
    MontgomeryInt out(*this);
    return out.MulConstantInPlace(c, cb, params);
  }
  // montgomery.h:383-415
  MontgomeryInt& AddInPlace(const MontgomeryInt& that, const Params* params) {
    n_ = params->BarrettReduce(n_ + that.n_);
    return *this;
  }
  MontgomeryInt Add(const MontgomeryInt& that, const Params* params) const {
    MontgomeryInt out(*this);
    return out.AddInPlace(that, params);
  }
  MontgomeryInt& NegateInPlace(const Params* params) {
    n_ = params->modulus - n_;
    return *this;
  }
  MontgomeryInt Negate(const Params* params) const { return MontgomeryInt(params->modulus - n_); }
  MontgomeryInt& SubInPlace(const MontgomeryInt& that, const Params* params) {
    n_ = params->BarrettReduce(n_ + (params->modulus - that.n_));
    return *this;
  }
  MontgomeryInt Sub(const MontgomeryInt& that, const Params* params) const {
    MontgomeryInt out(*this);
    return out.SubInPlace(that, params);
  }
  MontgomeryInt& LazyAddInPlace(const MontgomeryInt& that, const Params*) {
    n_ += that.n_;
    return *this;
  }
  MontgomeryInt& LazySubInPlace(const MontgomeryInt& that, const Params* params) {
    n_ += (params->modulus - that.n_);
    return *this;
  }
  // montgomery.h:446-465
  MontgomeryInt& FusedMulAddInPlace(const MontgomeryInt& a, const MontgomeryInt& b,
                                    const Params* params) {
    BigInt t = static_cast<BigInt>(params->r_mod_modulus) * n_ + static_cast<BigInt>(a.n_) * b.n_;
    Int u = static_cast<Int>(t) * params->inv_modulus;
    t += params->modulus_bigint * u;
    Int t_msb = static_cast<Int>(t >> 64);
    n_ = (t_msb >= params->modulus) ? t_msb - params->modulus : t_msb;
    return *this;
  }
  // montgomery.cc:473-499
  MontgomeryInt ModExp(Int exponent, const Params* params) const {
    MontgomeryInt result = ImportOne(params);
    MontgomeryInt base = *this;
    while (exponent > 0) {
      if (exponent % 2 == 1) result.MulInPlace(base, params);
      base.MulInPlace(base, params);
      exponent >>= 1;
    }
    return result;
  }
  MontgomeryInt MultiplicativeInverse(const Params* params) const {
    return ModExp(params->modulus - 2, params);
  }

  bool operator==(const MontgomeryInt& that) const { return n_ == that.n_; }
  bool operator!=(const MontgomeryInt& that) const { return n_ != that.n_; }

 private:
  Int n_;
};

using ModularInt64 = MontgomeryInt<Uint64>;

}  // namespace rlwe

#endif  // SHG_RLWE_MONTGOMERY_H_
