// Prints outputs of the host PRNG primitives (rlwe/prng.h) for the CPU test
// tests/test_cpp_prng.py, which checks them against Python's hashlib/hmac
// and the ChaCha20 keystreams in tests/golden/chacha_vectors.json.
//
//   prng_host_main sha256 <hex>           -> hex digest
//   prng_host_main hmac <hexkey> <hexmsg> -> hex tag
//   prng_host_main hkdf <hexikm> <salt> <len>  -> hex okm (empty info)
//   prng_host_main chacha <hexkey> <nbytes>    -> hex of the first nbytes Rand8 draws
//   prng_host_main hkdfprng <hexkey> <nbytes>  -> same for SingleThreadHkdfPrng
//   prng_host_main rand64 <hexkey> <count>     -> Rand64 draws of the ChaCha PRNG, decimal
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "rlwe/prng.h"

static std::string Unhex(const std::string& h) {
  std::string s;
  for (size_t i = 0; i + 1 < h.size(); i += 2) s.push_back((char)std::stoi(h.substr(i, 2), nullptr, 16));
  return s;
}
static void PrintHex(const uint8_t* p, size_t n) {
  for (size_t i = 0; i < n; ++i) std::printf("%02x", p[i]);
  std::printf("\n");
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string cmd = argv[1];
  if (cmd == "sha256") {
    const std::string m = Unhex(argv[2]);
    uint8_t d[32];
    rlwe::internal::Sha256((const uint8_t*)m.data(), m.size(), d);
    PrintHex(d, 32);
  } else if (cmd == "hmac" && argc >= 4) {
    uint8_t d[32];
    rlwe::internal::HmacSha256(Unhex(argv[2]), Unhex(argv[3]), d);
    PrintHex(d, 32);
  } else if (cmd == "hkdf" && argc >= 5) {
    auto okm = rlwe::internal::HkdfSha256(Unhex(argv[2]), argv[3], "", std::atoi(argv[4]));
    PrintHex(okm.data(), okm.size());
  } else if ((cmd == "chacha" || cmd == "hkdfprng") && argc >= 4) {
    std::unique_ptr<rlwe::SecurePrng> g;
    if (cmd == "chacha") {
      auto p = rlwe::SingleThreadChaChaPrng::Create(Unhex(argv[2]));
      if (!p.ok()) { std::printf("error %s\n", p.status().message().c_str()); return 1; }
      g = std::move(p).value();
    } else {
      auto p = rlwe::SingleThreadHkdfPrng::Create(Unhex(argv[2]));
      if (!p.ok()) { std::printf("error %s\n", p.status().message().c_str()); return 1; }
      g = std::move(p).value();
    }
    const int n = std::atoi(argv[3]);
    std::vector<uint8_t> out;
    for (int i = 0; i < n; ++i) out.push_back(g->Rand8().value());
    PrintHex(out.data(), out.size());
  } else if (cmd == "rand64" && argc >= 4) {
    auto g = rlwe::SingleThreadChaChaPrng::Create(Unhex(argv[2])).value();
    for (int i = 0; i < std::atoi(argv[3]); ++i)
      std::printf("%llu\n", (unsigned long long)g->Rand64().value());
  } else {
    return 2;
  }
  return 0;
}
