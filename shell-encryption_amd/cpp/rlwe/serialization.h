// The serialized messages of the RNS path (rns/rns_serialization.proto:27-67)
// and PrngType (serialization.proto:45-49), with the protobuf binary
// encoding written out by hand (no protobuf runtime is linked):
//   SerializedRnsPolynomial     1: int32 log_n   2: repeated bytes coeff_vectors
//                               3: bool is_ntt
//   SerializedRnsRlweCiphertext 1: repeated SerializedRnsPolynomial components
//                               2: int32 power_of_s   3: double error
//   SerializedRnsGaloisKey      2: repeated SerializedRnsPolynomial key_bs
//                               3: int32 power   4: bytes prng_seed   5: PrngType
// SerializeAsString writes the set fields in field-number order, as protobuf
// does, so the bytes are those a protobuf build of the reference would write;
// ParseFromString accepts fields in any order and skips unknown ones.
#ifndef SHG_RLWE_SERIALIZATION_H_
#define SHG_RLWE_SERIALIZATION_H_

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace rlwe {

enum PrngType { PRNG_TYPE_INVALID = 0, PRNG_TYPE_HKDF = 1, PRNG_TYPE_CHACHA = 2 };

namespace wire {

inline void PutVarint(std::string* out, uint64_t v) {
  while (v >= 0x80) {
    out->push_back((char)(v | 0x80));
    v >>= 7;
  }
  out->push_back((char)v);
}
inline void PutTag(std::string* out, int field, int wire_type) {
  PutVarint(out, ((uint64_t)field << 3) | (uint64_t)wire_type);
}
// int32 and enum fields are varints of the sign-extended 64-bit value.
inline void PutInt32(std::string* out, int field, int32_t v) {
  PutTag(out, field, 0);
  PutVarint(out, (uint64_t)(int64_t)v);
}
inline void PutBytes(std::string* out, int field, const std::string& v) {
  PutTag(out, field, 2);
  PutVarint(out, v.size());
  out->append(v);
}
inline void PutDouble(std::string* out, int field, double v) {
  PutTag(out, field, 1);
  uint64_t bits;
  std::memcpy(&bits, &v, 8);
  for (int i = 0; i < 8; ++i) out->push_back((char)(bits >> (8 * i)));
}

// A cursor over an encoded message.
class Reader {
 public:
  explicit Reader(const std::string& s) : p_((const uint8_t*)s.data()), end_(p_ + s.size()) {}
  bool done() const { return p_ == end_; }
  bool Varint(uint64_t* v) {
    *v = 0;
    for (int shift = 0; shift < 64; shift += 7) {
      if (p_ == end_) return false;
      const uint8_t b = *p_++;
      *v |= (uint64_t)(b & 0x7f) << shift;
      if (!(b & 0x80)) return true;
    }
    return false;
  }
  bool Tag(int* field, int* wire_type) {
    uint64_t t;
    if (!Varint(&t)) return false;
    *field = (int)(t >> 3);
    *wire_type = (int)(t & 7);
    return *field > 0;
  }
  bool Bytes(std::string* v) {
    uint64_t n;
    if (!Varint(&n) || n > (uint64_t)(end_ - p_)) return false;
    v->assign((const char*)p_, (size_t)n);
    p_ += n;
    return true;
  }
  bool Fixed64(uint64_t* v) {
    if (end_ - p_ < 8) return false;
    *v = 0;
    for (int i = 0; i < 8; ++i) *v |= (uint64_t)p_[i] << (8 * i);
    p_ += 8;
    return true;
  }
  bool Skip(int wire_type) {
    uint64_t v;
    std::string s;
    switch (wire_type) {
      case 0: return Varint(&v);
      case 1: return Fixed64(&v);
      case 2: return Bytes(&s);
      case 5:
        if (end_ - p_ < 4) return false;
        p_ += 4;
        return true;
      default: return false;
    }
  }

 private:
  const uint8_t* p_;
  const uint8_t* end_;
};

}  // namespace wire

class SerializedRnsPolynomial {
 public:
  int32_t log_n() const { return log_n_; }
  void set_log_n(int32_t v) { log_n_ = v; has_log_n_ = true; }
  int coeff_vectors_size() const { return (int)coeff_vectors_.size(); }
  const std::string& coeff_vectors(int i) const { return coeff_vectors_[i]; }
  std::string* add_coeff_vectors() { coeff_vectors_.emplace_back(); return &coeff_vectors_.back(); }
  bool is_ntt() const { return is_ntt_; }
  void set_is_ntt(bool v) { is_ntt_ = v; has_is_ntt_ = true; }

  std::string SerializeAsString() const {
    std::string out;
    if (has_log_n_) wire::PutInt32(&out, 1, log_n_);
    for (const auto& v : coeff_vectors_) wire::PutBytes(&out, 2, v);
    if (has_is_ntt_) {
      wire::PutTag(&out, 3, 0);
      wire::PutVarint(&out, is_ntt_ ? 1 : 0);
    }
    return out;
  }
  bool ParseFromString(const std::string& s) {
    *this = SerializedRnsPolynomial();
    wire::Reader r(s);
    while (!r.done()) {
      int f, wt;
      uint64_t v;
      if (!r.Tag(&f, &wt)) return false;
      if (f == 1 && wt == 0) {
        if (!r.Varint(&v)) return false;
        set_log_n((int32_t)v);
      } else if (f == 2 && wt == 2) {
        if (!r.Bytes(add_coeff_vectors())) return false;
      } else if (f == 3 && wt == 0) {
        if (!r.Varint(&v)) return false;
        set_is_ntt(v != 0);
      } else if (!r.Skip(wt)) {
        return false;
      }
    }
    return true;
  }

 private:
  int32_t log_n_ = 0;
  bool has_log_n_ = false;
  std::vector<std::string> coeff_vectors_;
  bool is_ntt_ = false;
  bool has_is_ntt_ = false;
};

class SerializedRnsRlweCiphertext {
 public:
  int components_size() const { return (int)components_.size(); }
  const SerializedRnsPolynomial& components(int i) const { return components_[i]; }
  SerializedRnsPolynomial* add_components() { components_.emplace_back(); return &components_.back(); }
  int32_t power_of_s() const { return power_of_s_; }
  void set_power_of_s(int32_t v) { power_of_s_ = v; has_power_ = true; }
  double error() const { return error_; }
  void set_error(double v) { error_ = v; has_error_ = true; }

  std::string SerializeAsString() const {
    std::string out;
    for (const auto& c : components_) wire::PutBytes(&out, 1, c.SerializeAsString());
    if (has_power_) wire::PutInt32(&out, 2, power_of_s_);
    if (has_error_) wire::PutDouble(&out, 3, error_);
    return out;
  }
  bool ParseFromString(const std::string& s) {
    *this = SerializedRnsRlweCiphertext();
    wire::Reader r(s);
    while (!r.done()) {
      int f, wt;
      uint64_t v;
      std::string b;
      if (!r.Tag(&f, &wt)) return false;
      if (f == 1 && wt == 2) {
        if (!r.Bytes(&b) || !add_components()->ParseFromString(b)) return false;
      } else if (f == 2 && wt == 0) {
        if (!r.Varint(&v)) return false;
        set_power_of_s((int32_t)v);
      } else if (f == 3 && wt == 1) {
        if (!r.Fixed64(&v)) return false;
        double d;
        std::memcpy(&d, &v, 8);
        set_error(d);
      } else if (!r.Skip(wt)) {
        return false;
      }
    }
    return true;
  }

 private:
  std::vector<SerializedRnsPolynomial> components_;
  int32_t power_of_s_ = 0;
  bool has_power_ = false;
  double error_ = 0;
  bool has_error_ = false;
};

class SerializedRnsGaloisKey {
 public:
  int key_bs_size() const { return (int)key_bs_.size(); }
  const SerializedRnsPolynomial& key_bs(int i) const { return key_bs_[i]; }
  SerializedRnsPolynomial* add_key_bs() { key_bs_.emplace_back(); return &key_bs_.back(); }
  int32_t power() const { return power_; }
  void set_power(int32_t v) { power_ = v; has_power_ = true; }
  const std::string& prng_seed() const { return prng_seed_; }
  void set_prng_seed(const std::string& v) { prng_seed_ = v; has_seed_ = true; }
  PrngType prng_type() const { return prng_type_; }
  void set_prng_type(PrngType v) { prng_type_ = v; has_type_ = true; }

  std::string SerializeAsString() const {
    std::string out;
    for (const auto& k : key_bs_) wire::PutBytes(&out, 2, k.SerializeAsString());
    if (has_power_) wire::PutInt32(&out, 3, power_);
    if (has_seed_) wire::PutBytes(&out, 4, prng_seed_);
    if (has_type_) wire::PutInt32(&out, 5, (int32_t)prng_type_);
    return out;
  }
  bool ParseFromString(const std::string& s) {
    *this = SerializedRnsGaloisKey();
    wire::Reader r(s);
    while (!r.done()) {
      int f, wt;
      uint64_t v;
      std::string b;
      if (!r.Tag(&f, &wt)) return false;
      if (f == 2 && wt == 2) {
        if (!r.Bytes(&b) || !add_key_bs()->ParseFromString(b)) return false;
      } else if (f == 3 && wt == 0) {
        if (!r.Varint(&v)) return false;
        set_power((int32_t)v);
      } else if (f == 4 && wt == 2) {
        if (!r.Bytes(&b)) return false;
        set_prng_seed(b);
      } else if (f == 5 && wt == 0) {
        if (!r.Varint(&v)) return false;
        set_prng_type((PrngType)(int32_t)v);
      } else if (!r.Skip(wt)) {
        return false;
      }
    }
    return true;
  }

 private:
  std::vector<SerializedRnsPolynomial> key_bs_;
  int32_t power_ = 0;
  bool has_power_ = false;
  std::string prng_seed_;
  bool has_seed_ = false;
  PrngType prng_type_ = PRNG_TYPE_INVALID;
  bool has_type_ = false;
};

}  // namespace rlwe

#endif  // SHG_RLWE_SERIALIZATION_H_
