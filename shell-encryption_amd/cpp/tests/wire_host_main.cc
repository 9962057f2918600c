// Encodes and re-encodes the serialized messages of rlwe/serialization.h for
// the CPU test tests/test_cpp_wire.py, which compares the bytes with the
// protobuf runtime's encoding of the same messages.
//
//   wire_host_main poly <log_n> <is_ntt> [hex coeff vector ...]
//   wire_host_main ct <power_of_s> <error as 16 hex digits of its bits> [hex poly msg ...]
//   wire_host_main gk <power> <hex seed> <prng type> [hex poly msg ...]
//   wire_host_main reparse <poly|ct|gk> <hex msg>
// Each prints the hex encoding of the message (or "parse error").
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "rlwe/serialization.h"

static std::string Unhex(const std::string& h) {
  std::string s;
  for (size_t i = 0; i + 1 < h.size(); i += 2) s.push_back((char)std::stoi(h.substr(i, 2), nullptr, 16));
  return s;
}
static void PrintHex(const std::string& s) {
  for (unsigned char c : s) std::printf("%02x", c);
  std::printf("\n");
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string cmd = argv[1];
  if (cmd == "poly" && argc >= 4) {
    rlwe::SerializedRnsPolynomial p;
    p.set_log_n(std::atoi(argv[2]));
    for (int i = 4; i < argc; ++i) *p.add_coeff_vectors() = Unhex(argv[i]);
    p.set_is_ntt(std::atoi(argv[3]) != 0);
    PrintHex(p.SerializeAsString());
  } else if (cmd == "ct" && argc >= 4) {
    rlwe::SerializedRnsRlweCiphertext c;
    for (int i = 4; i < argc; ++i)
      if (!c.add_components()->ParseFromString(Unhex(argv[i]))) { std::printf("parse error\n"); return 1; }
    c.set_power_of_s(std::atoi(argv[2]));
    const unsigned long long bits = std::strtoull(argv[3], nullptr, 16);
    double e;
    std::memcpy(&e, &bits, 8);
    c.set_error(e);
    PrintHex(c.SerializeAsString());
  } else if (cmd == "gk" && argc >= 5) {
    rlwe::SerializedRnsGaloisKey g;
    for (int i = 5; i < argc; ++i)
      if (!g.add_key_bs()->ParseFromString(Unhex(argv[i]))) { std::printf("parse error\n"); return 1; }
    g.set_power(std::atoi(argv[2]));
    g.set_prng_seed(Unhex(argv[3]));
    g.set_prng_type((rlwe::PrngType)std::atoi(argv[4]));
    PrintHex(g.SerializeAsString());
  } else if (cmd == "reparse" && argc >= 4) {
    const std::string kind = argv[2], msg = Unhex(argv[3]);
    std::string out;
    bool ok = false;
    if (kind == "poly") { rlwe::SerializedRnsPolynomial m; ok = m.ParseFromString(msg); out = m.SerializeAsString(); }
    if (kind == "ct") { rlwe::SerializedRnsRlweCiphertext m; ok = m.ParseFromString(msg); out = m.SerializeAsString(); }
    if (kind == "gk") { rlwe::SerializedRnsGaloisKey m; ok = m.ParseFromString(msg); out = m.SerializeAsString(); }
    if (!ok) { std::printf("parse error\n"); return 0; }
    PrintHex(out);
  } else {
    return 2;
  }
  return 0;
}
