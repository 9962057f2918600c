// SecurePrng (prng/prng.h:28-37) and the reference's two implementations,
// SingleThreadChaChaPrng (prng/single_thread_chacha_prng.cc) and
// SingleThreadHkdfPrng (prng/single_thread_hkdf_prng.cc), on the host.
//
// Both hand out bytes of a buffer of 255 * 32 bytes, drawn again ("resalted")
// when used up; Rand64 is eight Rand8 in little-endian order
// (prng/chacha_prng_util.cc:76-104, prng/hkdf_prng_util.cc:46-76):
//   ChaCha: buffer s = ChaCha20 keystream (RFC 7539, block counter from 0)
//           under the 32-byte key and nonce "salt" || LE64(s)
//           (prng/chacha_prng_util.cc:27-60);
//   HKDF:   buffer s = HKDF-SHA256 (RFC 5869) of the 64-byte key with salt
//           "salt" followed by s in decimal and empty info
//           (prng/hkdf_prng_util.cc:27-43).
// The reference takes these primitives from BoringSSL and Tink; they are
// written here from the RFCs and FIPS 180-4.  SHA-256's round constants and
// initial hash are derived at first use from their definitions (the first 32
// fractional bits of the cube roots of the first 64 primes and of the square
// roots of the first 8), with exact integer roots.
#ifndef SHG_RLWE_PRNG_H_
#define SHG_RLWE_PRNG_H_

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "rlwe/status.h"

namespace rlwe {

class SecurePrng {
 public:
  virtual StatusOr<uint8_t> Rand8() = 0;
  virtual StatusOr<uint64_t> Rand64() = 0;
  virtual ~SecurePrng() = default;
};

namespace internal {

constexpr int kPrngBufferBytes = 255 * 32;

inline uint32_t Le32(const uint8_t* p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
inline uint32_t Rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
inline uint32_t Rotr32(uint32_t v, int c) { return (v >> c) | (v << (32 - c)); }

// ChaCha20 block function (RFC 7539 Sec. 2.3): 64 keystream bytes.
inline void ChaCha20Block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                          uint8_t out[64]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
  for (int i = 0; i < 8; ++i) s[4 + i] = Le32(key + 4 * i);
  s[12] = counter;
  for (int i = 0; i < 3; ++i) s[13 + i] = Le32(nonce + 4 * i);
  uint32_t x[16];
  std::memcpy(x, s, sizeof(s));
  auto qr = [&x](int a, int b, int c, int d) {
    x[a] += x[b]; x[d] = Rotl32(x[d] ^ x[a], 16);
    x[c] += x[d]; x[b] = Rotl32(x[b] ^ x[c], 12);
    x[a] += x[b]; x[d] = Rotl32(x[d] ^ x[a], 8);
    x[c] += x[d]; x[b] = Rotl32(x[b] ^ x[c], 7);
  };
  for (int r = 0; r < 10; ++r) {
    qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
    qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
  }
  for (int i = 0; i < 16; ++i) {
    const uint32_t v = x[i] + s[i];
    for (int b = 0; b < 4; ++b) out[4 * i + b] = (uint8_t)(v >> (8 * b));
  }
}

// floor(v^(1/k)) for k = 2 or 3 and v < 2^120, by bisection.
inline uint64_t IntegerRoot(unsigned __int128 v, int k) {
  uint64_t lo = 0, hi = (uint64_t)1 << (k == 2 ? 60 : 40);
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    unsigned __int128 p = (unsigned __int128)mid * mid;
    if (k == 3) p *= mid;
    if (p <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct Sha256Constants {
  uint32_t k[64];
  uint32_t h0[8];
  Sha256Constants() {
    int found = 0;
    for (uint64_t p = 2; found < 64; ++p) {
      bool prime = true;
      for (uint64_t d = 2; d * d <= p; ++d)
        if (p % d == 0) { prime = false; break; }
      if (!prime) continue;
      // Fractional bits of cbrt(p): floor(cbrt(p * 2^96)) mod 2^32.
      k[found] = (uint32_t)IntegerRoot((unsigned __int128)p << 96, 3);
      if (found < 8) h0[found] = (uint32_t)IntegerRoot((unsigned __int128)p << 64, 2);
      ++found;
    }
  }
};
inline const Sha256Constants& Sha256K() {
  static const Sha256Constants c;
  return c;
}

// SHA-256 (FIPS 180-4).
inline void Sha256(const uint8_t* data, size_t len, uint8_t out[32]) {
  const Sha256Constants& C = Sha256K();
  uint32_t h[8];
  std::memcpy(h, C.h0, sizeof(h));
  std::vector<uint8_t> msg(data, data + len);
  msg.push_back(0x80);
  while (msg.size() % 64 != 56) msg.push_back(0);
  const uint64_t bits = (uint64_t)len * 8;
  for (int i = 7; i >= 0; --i) msg.push_back((uint8_t)(bits >> (8 * i)));
  for (size_t blk = 0; blk < msg.size(); blk += 64) {
    uint32_t w[64];
    for (int t = 0; t < 16; ++t) {
      const uint8_t* q = &msg[blk + 4 * t];
      w[t] = (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3];
    }
    for (int t = 16; t < 64; ++t) {
      const uint32_t s0 = Rotr32(w[t - 15], 7) ^ Rotr32(w[t - 15], 18) ^ (w[t - 15] >> 3);
      const uint32_t s1 = Rotr32(w[t - 2], 17) ^ Rotr32(w[t - 2], 19) ^ (w[t - 2] >> 10);
      w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int t = 0; t < 64; ++t) {
      const uint32_t S1 = Rotr32(e, 6) ^ Rotr32(e, 11) ^ Rotr32(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + C.k[t] + w[t];
      const uint32_t S0 = Rotr32(a, 2) ^ Rotr32(a, 13) ^ Rotr32(a, 22);
      const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
      const uint32_t t2 = S0 + maj;
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }
  for (int i = 0; i < 8; ++i)
    for (int b = 0; b < 4; ++b) out[4 * i + b] = (uint8_t)(h[i] >> (24 - 8 * b));
}

// HMAC-SHA256 (RFC 2104).
inline void HmacSha256(const std::string& key, const std::string& msg, uint8_t out[32]) {
  uint8_t k[64] = {0};
  if (key.size() > 64) {
    Sha256((const uint8_t*)key.data(), key.size(), k);
  } else {
    std::memcpy(k, key.data(), key.size());
  }
  std::string inner(64, '\0'), outer(64, '\0');
  for (int i = 0; i < 64; ++i) {
    inner[i] = (char)(k[i] ^ 0x36);
    outer[i] = (char)(k[i] ^ 0x5c);
  }
  inner += msg;
  uint8_t ih[32];
  Sha256((const uint8_t*)inner.data(), inner.size(), ih);
  outer.append((const char*)ih, 32);
  Sha256((const uint8_t*)outer.data(), outer.size(), out);
}

// HKDF-SHA256 (RFC 5869): extract with `salt`, expand with `info`.
inline std::vector<uint8_t> HkdfSha256(const std::string& ikm, const std::string& salt,
                                       const std::string& info, size_t length) {
  uint8_t prk[32];
  HmacSha256(salt, ikm, prk);
  const std::string prk_s((const char*)prk, 32);
  std::vector<uint8_t> okm;
  std::string t;
  for (uint8_t i = 1; okm.size() < length; ++i) {
    uint8_t ti[32];
    HmacSha256(prk_s, t + info + std::string(1, (char)i), ti);
    t.assign((const char*)ti, 32);
    for (int b = 0; b < 32 && okm.size() < length; ++b) okm.push_back(ti[b]);
  }
  return okm;
}

inline StatusOr<std::string> RandomSeed(size_t bytes) {
  std::random_device rd;  // the platform's entropy source
  std::string s(bytes, '\0');
  for (size_t i = 0; i < bytes; ++i) s[i] = (char)(rd() & 0xff);
  return s;
}

// Buffer-based PRNG of the reference: `Refill(salt)` returns buffer `salt`.
class BufferedPrng : public SecurePrng {
 public:
  StatusOr<uint8_t> Rand8() override {
    if (position_ >= (int)buffer_.size()) Resalt();
    return buffer_[position_++];
  }
  StatusOr<uint64_t> Rand64() override {
    uint64_t r = 0;
    for (int i = 0; i < 8; ++i) {
      RLWE_ASSIGN_OR_RETURN(uint8_t b, Rand8());
      r += (uint64_t)b << (8 * i);
    }
    return r;
  }

 protected:
  explicit BufferedPrng(std::string key) : key_(std::move(key)) {}
  virtual std::vector<uint8_t> Refill(int salt) const = 0;
  void Resalt() {
    buffer_ = Refill(salt_counter_);
    ++salt_counter_;
    position_ = 0;
  }
  std::string key_;

 private:
  int position_ = 0;
  int salt_counter_ = 0;
  std::vector<uint8_t> buffer_;
};

}  // namespace internal

class SingleThreadChaChaPrng : public internal::BufferedPrng {
 public:
  static StatusOr<std::unique_ptr<SingleThreadChaChaPrng>> Create(const std::string& in_key) {
    if ((int)in_key.size() != SeedLength())
      return InvalidArgumentError("Cannot create Prng with key of the wrong size. Real key length of " +
                                  std::to_string(in_key.size()) +
                                  " instead of expected key length of " +
                                  std::to_string(SeedLength()) + ".");
    std::unique_ptr<SingleThreadChaChaPrng> p(new SingleThreadChaChaPrng(in_key));
    p->Resalt();
    return p;
  }
  static StatusOr<std::string> GenerateSeed() { return internal::RandomSeed(SeedLength()); }
  static int SeedLength() { return 32; }

 private:
  explicit SingleThreadChaChaPrng(std::string key) : BufferedPrng(std::move(key)) {}
  std::vector<uint8_t> Refill(int salt) const override {
    uint8_t nonce[12] = {'s', 'a', 'l', 't'};
    uint64_t c = (uint64_t)salt;
    for (int i = 0; i < 8; ++i, c >>= 8) nonce[4 + i] = (uint8_t)(c & 0xff);
    std::vector<uint8_t> buf(internal::kPrngBufferBytes);
    for (int blk = 0; blk * 64 < internal::kPrngBufferBytes; ++blk) {
      uint8_t ks[64];
      internal::ChaCha20Block((const uint8_t*)key_.data(), (uint32_t)blk, nonce, ks);
      const int take = std::min(64, internal::kPrngBufferBytes - blk * 64);
      std::memcpy(buf.data() + blk * 64, ks, (size_t)take);
    }
    return buf;
  }
};

class SingleThreadHkdfPrng : public internal::BufferedPrng {
 public:
  static StatusOr<std::unique_ptr<SingleThreadHkdfPrng>> Create(const std::string& in_key) {
    if ((int)in_key.size() != SeedLength())
      return InvalidArgumentError("Cannot create Prng with key of the wrong size. Real key length of " +
                                  std::to_string(in_key.size()) +
                                  " instead of expected key length of " +
                                  std::to_string(SeedLength()) + ".");
    std::unique_ptr<SingleThreadHkdfPrng> p(new SingleThreadHkdfPrng(in_key));
    p->Resalt();
    return p;
  }
  static StatusOr<std::string> GenerateSeed() { return internal::RandomSeed(SeedLength()); }
  static int SeedLength() { return 64; }

 private:
  explicit SingleThreadHkdfPrng(std::string key) : BufferedPrng(std::move(key)) {}
  std::vector<uint8_t> Refill(int salt) const override {
    return internal::HkdfSha256(key_, "salt" + std::to_string(salt), "",
                                internal::kPrngBufferBytes);
  }
};

}  // namespace rlwe

#endif  // SHG_RLWE_PRNG_H_
