// GPU tests of the rlwe C++ API (the drop-in boundary): each case mirrors a
// reference test (cited) and checks the device results bit for bit against
// the CPU oracle (oracle/shell_oracle.c, linked into this test binary only).
//
// Run: rns_api_test  (needs an MI355X; exits non-zero on the first failure)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "rlwe/rns_context.h"
#include "rlwe/rns_relinearization_key.h"
#include "rlwe/rns_secret_key.h"
#include "shell_oracle.h"

using ModularInt = rlwe::MontgomeryInt<uint64_t>;
using Poly = rlwe::RnsPolynomial<ModularInt>;
using Ct = rlwe::RnsBgvCiphertext<ModularInt>;

static int g_checks = 0;
#define CHECK(cond)                                                            \
  do {                                                                         \
    ++g_checks;                                                                \
    if (!(cond)) {                                                             \
      std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)
#define CHECK_OK(expr)                                                          \
  do {                                                                          \
    auto _st = (expr);                                                          \
    if (!_st.ok()) {                                                            \
      std::fprintf(stderr, "%s:%d: not ok: %s\n", __FILE__, __LINE__, _st.message().c_str()); \
      std::exit(1);                                                             \
    }                                                                           \
  } while (0)
#define ASSIGN_OR_DIE(lhs, expr)                                                 \
  auto lhs##_or = (expr);                                                        \
  if (!lhs##_or.ok()) {                                                          \
    std::fprintf(stderr, "%s:%d: not ok: %s\n", __FILE__, __LINE__,              \
                 lhs##_or.status().message().c_str());                          \
    std::exit(1);                                                                \
  }                                                                              \
  auto lhs = std::move(lhs##_or).value()

static bool IsPrime(uint64_t n) {
  if (n < 2) return false;
  for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull})
    if (n % p == 0) return n == p;
  uint64_t d = n - 1;
  int s = 0;
  while (d % 2 == 0) { d /= 2; ++s; }
  auto mulmod = [n](uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % n); };
  for (uint64_t a : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
    uint64_t x = 1, b = a, e = d;
    while (e) { if (e & 1) x = mulmod(x, b); b = mulmod(b, b); e >>= 1; }
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; ++r) { x = mulmod(x, x); if (x == n - 1) { comp = false; break; } }
    if (comp) return false;
  }
  return true;
}

// The `count` largest primes below 2^log_bound with q = 1 mod 2^(log_n+1).
static std::vector<uint64_t> NttPrimes(int log_n, int count, int log_bound = 50) {
  const uint64_t step = 1ull << (log_n + 1);
  uint64_t q = ((1ull << log_bound) - 1) / step * step + 1;
  std::vector<uint64_t> out;
  while ((int)out.size() < count) {
    if (IsPrime(q)) out.push_back(q);
    q -= step;
  }
  return out;
}

struct OracleChain {
  std::vector<so_prime*> primes;
  explicit OracleChain(const std::vector<uint64_t>& qs, int log_n) {
    for (uint64_t q : qs) {
      int st = 0;
      primes.push_back(so_prime_create(q, log_n, &st));
    }
  }
  ~OracleChain() {
    for (auto* p : primes) so_prime_destroy(p);
  }
  so_prime* const* pp() const { return primes.data(); }
};

static std::vector<uint64_t> Flat(const Poly& p) {
  std::vector<uint64_t> out;
  for (const auto& v : p.Coeffs())
    for (const auto& x : v) out.push_back(x.GetMontgomeryRepresentation());
  return out;
}

static Poly FromFlat(const std::vector<uint64_t>& flat, int L, int n, bool ntt) {
  std::vector<std::vector<ModularInt>> cv(L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < n; ++j) cv[i].emplace_back(flat[(size_t)i * n + j]);
  return Poly::Create(std::move(cv), ntt).value();
}

static std::vector<uint64_t> Uniform(const OracleChain& oc, int L, uint64_t seed) {
  const size_t n = (size_t)1 << oc.primes[0]->log_n;
  std::vector<uint64_t> x((size_t)L * n);
  so_fill_uniform(oc.pp(), L, seed, x.data());
  return x;
}

// montgomery_test.cc:342-534 (host MontgomeryInt against the oracle).
static void TestMontgomery() {
  for (uint64_t q : NttPrimes(12, 2)) {
    auto params = rlwe::MontgomeryIntParams<uint64_t>::Create(q).value();
    so_mont m;
    so_mont_init(q, &m);
    CHECK(params->inv_modulus == m.inv_modulus && params->inv_r == m.inv_r);
    std::mt19937_64 rng(q);
    for (int i = 0; i < 1000; ++i) {
      uint64_t a = rng() % q, b = rng() % q;
      ModularInt x(a), y(b);
      CHECK(x.Mul(y, params.get()).GetMontgomeryRepresentation() == so_mul(&m, a, b));
      CHECK(x.Add(y, params.get()).GetMontgomeryRepresentation() == so_add(&m, a, b));
      CHECK(ModularInt::ImportInt(a, params.get()).value().GetMontgomeryRepresentation() ==
            so_import(&m, a));
    }
  }
  CHECK(!rlwe::MontgomeryIntParams<uint64_t>::Create(1ull << 62).ok());
}

// ntt_parameters_test.cc:71-286 (tables against the oracle).
static void TestNttParameters() {
  const int log_n = 10;
  for (uint64_t q : NttPrimes(log_n, 2)) {
    auto params = rlwe::MontgomeryIntParams<uint64_t>::Create(q).value();
    auto ntt = rlwe::InitializeNttParameters<ModularInt>(log_n, params.get()).value();
    int st = 0;
    so_prime* pr = so_prime_create(q, log_n, &st);
    for (int i = 0; i < (1 << log_n); ++i) {
      CHECK(std::get<0>(ntt.psis_bitrev_constant[i]) == pr->psis_bitrev_c[i]);
      CHECK(std::get<1>(ntt.psis_bitrev_constant[i]) == pr->psis_bitrev_cb[i]);
      CHECK(std::get<0>(ntt.psis_inv_bitrev_constant[i]) == pr->psis_inv_bitrev_c[i]);
      CHECK(ntt.bitrevs[i] == pr->bitrevs[i]);
    }
    CHECK(ntt.n_inv_ptr->GetMontgomeryRepresentation() == pr->n_inv);
    so_prime_destroy(pr);
  }
  auto params = rlwe::MontgomeryIntParams<uint64_t>::Create(97).value();
  CHECK(!rlwe::InitializeNttParameters<ModularInt>(6, params.get()).ok());  // 97 != 1 mod 128
  CHECK(!rlwe::InitializeNttParameters<ModularInt>(16, params.get()).ok());
}

// rns/rns_polynomial_test.cc:475-557,920-1109
static void TestRnsPolynomial() {
  const int log_n = 11, L = 3, n = 1 << log_n;
  auto qs = NttPrimes(log_n, L);
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, 65537).value();
  auto moduli = ctx.MainPrimeModuli();
  OracleChain oc(qs, log_n);
  auto x = Uniform(oc, L, 5), y = Uniform(oc, L, 6), z = Uniform(oc, L, 7);
  Poly p = FromFlat(x, L, n, false);
  // NTT / INTT against the oracle, and the round trip.
  CHECK_OK(p.ConvertToNttForm(moduli));
  auto ex = x;
  so_poly_ntt_fwd(oc.pp(), L, ex.data());
  CHECK(Flat(p) == ex);
  auto st = p.ConvertToNttForm(moduli);
  CHECK(!st.ok() && st.message() == "Polynomial already in NTT form");
  CHECK(Flat(p) == ex);  // unchanged on failure
  std::vector<const rlwe::PrimeModulus<ModularInt>*> two(moduli.begin(), moduli.begin() + 2);
  st = p.NegateInPlace(two);
  CHECK(!st.ok() && st.message() == "`moduli` must contain 3 RNS moduli.");
  CHECK_OK(p.ConvertToCoeffForm(moduli));
  CHECK(Flat(p) == x);
  CHECK_OK(p.ConvertToNttForm(moduli));
  // Pointwise operations.
  Poly a = FromFlat(y, L, n, true), b = FromFlat(z, L, n, true);
  Poly s = p.Add(a, moduli).value();
  auto es = ex;
  so_poly_pointwise(oc.pp(), L, SO_ADD, es.data(), y.data(), nullptr);
  CHECK(Flat(s) == es);
  CHECK_OK(s.FusedMulAddInPlace(a, b, moduli));
  so_poly_pointwise(oc.pp(), L, SO_FMA, es.data(), y.data(), z.data());
  CHECK(Flat(s) == es);
  CHECK_OK(s.MulInPlace(a, moduli));
  so_poly_pointwise(oc.pp(), L, SO_MUL, es.data(), y.data(), nullptr);
  CHECK(Flat(s) == es);
  CHECK_OK(s.MulInPlace(uint64_t{12345}, moduli));
  so_poly_mul_int(oc.pp(), L, es.data(), 12345);
  CHECK(Flat(s) == es);
  CHECK_OK(s.NegateInPlace(moduli));
  so_poly_pointwise(oc.pp(), L, SO_NEG, es.data(), nullptr, nullptr);
  CHECK(Flat(s) == es);
  Poly zero = Poly::CreateZero(log_n, moduli).value();
  CHECK_OK(zero.NegateInPlace(moduli));
  CHECK(zero.Coeffs()[0][0].GetMontgomeryRepresentation() == qs[0]);  // Negate(0) = q
  // a * 1 = a
  Poly one = Poly::CreateOne(log_n, moduli).value();
  CHECK(a.Mul(one, moduli).value() == a);
  // Substitute: power checks and values.
  CHECK(!p.Substitute(2, moduli).ok());
  CHECK(!p.Substitute(2 * n + 1, moduli).ok());
  CHECK(p.Substitute(1, moduli).value() == p);
  Poly sub = p.Substitute(5, moduli).value();
  std::vector<uint64_t> esub(ex.size());
  CHECK(so_poly_substitute(oc.pp(), L, 5, ex.data(), esub.data()) == 0);
  CHECK(Flat(sub) == esub);
  Poly coeff = FromFlat(x, L, n, false);
  CHECK(!coeff.Substitute(5, moduli).ok());
  CHECK(!coeff.MulInPlace(a, moduli).ok());
  std::printf("  RnsPolynomial: ok\n");
}

// rns/rns_polynomial_test.cc:1150-1332, rns/rns_gadget_test.cc:174-242
static void TestModReduceAndGadget() {
  const int log_n = 10, L = 4, n = 1 << log_n;
  auto qs = NttPrimes(log_n, L);
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, 65537).value();
  auto moduli = ctx.MainPrimeModuli();
  OracleChain oc(qs, log_n);
  auto x = Uniform(oc, L, 11);
  for (bool ntt : {true, false}) {
    Poly p = FromFlat(x, L, n, ntt);
    auto ql_inv = ctx.MainPrimeModulusInverseResidues()[L - 1].Prefix(L - 1);
    CHECK_OK(p.ModReduceLsb(65537, ql_inv, moduli));
    CHECK(p.NumModuli() == L - 1);
    std::vector<uint64_t> qi;
    for (auto& v : ql_inv.zs) qi.push_back(v.GetMontgomeryRepresentation());
    auto ex = x;
    CHECK(so_poly_mod_reduce_lsb(oc.pp(), L, 65537, qi.data(), ntt ? 1 : 0, ex.data()) == 0);
    ex.resize((size_t)(L - 1) * n);
    CHECK(Flat(p) == ex);
  }
  std::vector<size_t> log_bs(L, 10);
  auto q_hats = ctx.MainPrimeModulusComplements(L - 1).value();
  auto q_hat_invs = ctx.MainPrimeModulusCrtFactors(L - 1).value();
  auto gadget = rlwe::RnsGadget<ModularInt>::Create(log_n, log_bs, q_hats, q_hat_invs, moduli).value();
  CHECK(gadget.Dimension() == 20);
  Poly a = FromFlat(x, L, n, false);
  auto digits = gadget.Decompose(a, moduli).value();
  std::vector<uint64_t> ed((size_t)gadget.Dimension() * L * n);
  int lb[4] = {10, 10, 10, 10};
  so_gadget_decompose(oc.pp(), L, lb, x.data(), ed.data());
  for (int j = 0; j < gadget.Dimension(); ++j) {
    auto f = Flat(digits[j]);
    CHECK(std::equal(f.begin(), f.end(), ed.begin() + (size_t)j * L * n));
  }
  // <g, g^-1(a)> = a
  Poly acc = Poly::CreateZero(log_n, moduli, false).value();
  for (int j = 0; j < gadget.Dimension(); ++j) {
    Poly t = digits[j];
    CHECK_OK(t.MulInPlace(gadget.Component(j), moduli));
    CHECK_OK(acc.AddInPlace(t, moduli));
  }
  CHECK(Flat(acc) == x);
  std::printf("  ModReduceLsb and RnsGadget: ok\n");
}

// rns/rns_relinearization_key_test.cc:424-500, rns/rns_galois_key_test.cc:401-462,
// rns/rns_bgv_ciphertext_test.cc:293-329,660-713, given the key matrices.
static void TestKeySwitching(int log_n, int L, int batch) {
  const int n = 1 << log_n;
  auto qs = NttPrimes(log_n, L);
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, 65537).value();
  auto moduli = ctx.MainPrimeModuli();
  OracleChain oc(qs, log_n);
  std::vector<size_t> log_bs(L, 10);
  auto gadget = rlwe::RnsGadget<ModularInt>::Create(
      log_n, log_bs, ctx.MainPrimeModulusComplements(L - 1).value(),
      ctx.MainPrimeModulusCrtFactors(L - 1).value(), moduli).value();
  const int k = gadget.Dimension();
  auto error_params = rlwe::RnsErrorParams<ModularInt>::Create(log_n, moduli, {}, 17, 4.0).value();
  std::vector<uint64_t> kas((size_t)k * L * n), kbs((size_t)k * L * n);
  std::vector<Poly> pa, pb;
  for (int j = 0; j < k; ++j) {
    auto va = Uniform(oc, L, 100 + j), vb = Uniform(oc, L, 200 + j);
    std::copy(va.begin(), va.end(), kas.begin() + (size_t)j * L * n);
    std::copy(vb.begin(), vb.end(), kbs.begin() + (size_t)j * L * n);
    pa.push_back(FromFlat(va, L, n, true));
    pb.push_back(FromFlat(vb, L, n, true));
  }
  auto relin = rlwe::RnsRelinKey<ModularInt>::CreateFromKeyComponents(pa, pb, 2, &gadget, moduli).value();
  auto galois = rlwe::RnsGaloisKey<ModularInt>::CreateFromKeyComponents(pa, pb, 5, &gadget, moduli).value();
  std::vector<Ct> cts;
  std::vector<std::vector<uint64_t>> raw;
  for (int c = 0; c < batch; ++c) {
    std::vector<Poly> comps;
    std::vector<uint64_t> all;
    for (int i = 0; i < 2; ++i) {
      auto v = Uniform(oc, L, 1000 + 10 * c + i);
      all.insert(all.end(), v.begin(), v.end());
      comps.push_back(FromFlat(v, L, n, true));
    }
    raw.push_back(all);
    cts.emplace_back(std::move(comps), moduli, 1, 1.0, &error_params);
  }
  // Tensor product, then relinearization (single and batched).
  ASSIGN_OR_DIE(prod, cts[0] * cts[batch - 1]);
  CHECK(prod.Degree() == 2);
  std::vector<uint64_t> etp((size_t)3 * L * n);
  so_tensor_product(oc.pp(), L, 2, raw[0].data(), 2, raw[batch - 1].data(), etp.data());
  for (int i = 0; i < 3; ++i) {
    auto f = Flat(prod.Components()[i]);
    CHECK(std::equal(f.begin(), f.end(), etp.begin() + (size_t)i * L * n));
  }
  ASSIGN_OR_DIE(rel, relin.ApplyTo(prod));
  std::vector<uint64_t> e0((size_t)L * n), e1((size_t)L * n);
  int lb[64];
  for (int i = 0; i < L; ++i) lb[i] = 10;
  CHECK(so_relin_apply(oc.pp(), L, lb, 2, etp.data(), 7, kas.data(), kbs.data(), e0.data(), e1.data()) == 0);
  CHECK(Flat(rel.Components()[0]) == e0 && Flat(rel.Components()[1]) == e1);
  CHECK(rel.Degree() == 1 && rel.PowerOfS() == 1);
  CHECK(rel.Error() == prod.Error() + error_params.BoundOnGadgetBasedKeySwitching(1, 10, k));
  std::vector<Ct> prods;
  for (int c = 0; c < batch; ++c) prods.push_back((cts[c] * cts[(c + 1) % batch]).value());
  std::vector<const Ct*> ptrs;
  for (auto& c : prods) ptrs.push_back(&c);
  ASSIGN_OR_DIE(rels, relin.ApplyToBatch(ptrs));
  for (int c = 0; c < batch; ++c) {
    ASSIGN_OR_DIE(one, relin.ApplyTo(prods[c]));
    CHECK(one.Components()[0] == rels[c].Components()[0]);
    CHECK(one.Components()[1] == rels[c].Components()[1]);
  }
  // Galois: substitution then key switch.
  ASSIGN_OR_DIE(sub, cts[0].Substitute(5));
  CHECK(sub.PowerOfS() == 5);
  CHECK(!relin.ApplyTo(sub).ok());  // power of s must be 1
  ASSIGN_OR_DIE(gal, galois.ApplyTo(sub));
  std::vector<uint64_t> sraw((size_t)2 * L * n);
  so_poly_substitute(oc.pp(), L, 5, raw[0].data(), sraw.data());
  so_poly_substitute(oc.pp(), L, 5, raw[0].data() + (size_t)L * n, sraw.data() + (size_t)L * n);
  CHECK(so_galois_apply(oc.pp(), L, lb, sraw.data(), sraw.data() + (size_t)L * n, 1, kas.data(),
                        kbs.data(), e0.data(), e1.data()) == 0);
  CHECK(Flat(gal.Components()[0]) == e0 && Flat(gal.Components()[1]) == e1);
  CHECK(!galois.ApplyTo(cts[0]).ok());  // power mismatch
  // Modulus switch of the relinearized ciphertext.
  auto ql_inv = ctx.MainPrimeModulusInverseResidues()[L - 1].Prefix(L - 1);
  CHECK_OK(rel.ModReduce(65537, ql_inv));
  CHECK(rel.NumModuli() == L - 1);
  std::printf("  key switching at n=2^%d, L=%d, k=%d, batch %d: ok\n", log_n, L, k, batch);
}

// rns/rns_polynomial_test.cc (Serialize/Deserialize), rns/rns_ciphertext_test.cc:
// the bytes are the oracle's SerializeVector per limb, and the messages
// survive their binary encoding.
static void TestSerialization() {
  const int log_n = 12, L = 3, n = 1 << log_n;
  auto qs = NttPrimes(log_n, L);
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, 65537).value();
  auto moduli = ctx.MainPrimeModuli();
  OracleChain oc(qs, log_n);
  auto x = Uniform(oc, L, 77);
  x[5] = qs[0];  // NegateInPlace(0) = q fits the field width
  auto p = FromFlat(x, L, n, true);
  ASSIGN_OR_DIE(ser, p.Serialize(moduli));
  CHECK(ser.log_n() == log_n && ser.is_ntt() && ser.coeff_vectors_size() == L);
  for (int i = 0; i < L; ++i) {
    std::vector<uint8_t> ref(((size_t)n * 50 + 7) / 8 + 8);
    const long long nb = so_serialize_vector(&oc.primes[i]->mp, x.data() + (size_t)i * n, n,
                                             ref.data(), ref.size());
    CHECK(nb == (long long)ser.coeff_vectors(i).size());
    CHECK(std::memcmp(ref.data(), ser.coeff_vectors(i).data(), (size_t)nb) == 0);
  }
  rlwe::SerializedRnsPolynomial parsed;
  CHECK(parsed.ParseFromString(ser.SerializeAsString()));
  ASSIGN_OR_DIE(back, Poly::Deserialize(parsed, moduli));
  CHECK(back == p);
  // Reference error messages.
  auto bad = Poly::Deserialize(parsed, rlwe::Span<const rlwe::PrimeModulus<ModularInt>* const>(
                                           moduli.data(), 2));
  CHECK(!bad.ok() && bad.status().message().find("coefficient vectors must be 2") != std::string::npos);
  rlwe::SerializedRnsPolynomial shortv;
  shortv.set_log_n(log_n);
  for (int i = 0; i < L; ++i) *shortv.add_coeff_vectors() = ser.coeff_vectors(i).substr(0, 10);
  auto bad2 = Poly::Deserialize(shortv, moduli);
  CHECK(!bad2.ok() && bad2.status().message().find("too small") != std::string::npos);
  CHECK(!p.Serialize(rlwe::Span<const rlwe::PrimeModulus<ModularInt>* const>(moduli.data(), 2)).ok());
  // Ciphertext round trip through its encoding.
  auto error_params = rlwe::RnsErrorParams<ModularInt>::Create(log_n, moduli, {}, 17, 4.0).value();
  std::vector<Poly> comps = {p, FromFlat(Uniform(oc, L, 78), L, n, true)};
  Ct ct(std::move(comps), moduli, 5, 123.5, &error_params);
  ASSIGN_OR_DIE(sct, ct.Serialize());
  rlwe::SerializedRnsRlweCiphertext sct2;
  CHECK(sct2.ParseFromString(sct.SerializeAsString()));
  ASSIGN_OR_DIE(ct2, Ct::Deserialize(sct2, moduli, &error_params));
  CHECK(ct2.PowerOfS() == 5 && ct2.Error() == 123.5 && ct2.Len() == 2);
  CHECK(ct2.Components()[0] == ct.Components()[0] && ct2.Components()[1] == ct.Components()[1]);
  CHECK(!Ct::Deserialize(sct2, moduli, nullptr).ok());
  std::printf("  serialization: ok\n");
}

// The key relation of a fresh key column: b + a*s - target = t*e with e a
// centered binomial of the given variance, the same small integer on every
// limb (rns/rns_relinearization_key.cc:160-170).
static void CheckKeyColumn(const Poly& a, const Poly& b, const Poly& s, const Poly& target,
                           const std::vector<const rlwe::PrimeModulus<ModularInt>*>& moduli,
                           uint64_t t, int variance) {
  Poly r = b;
  CHECK_OK(r.FusedMulAddInPlace(a, s, moduli));
  CHECK_OK(r.SubInPlace(target, moduli));
  CHECK_OK(r.ConvertToCoeffForm(moduli));
  const auto& cv = r.Coeffs();
  const int n = r.NumCoeffs();
  for (int j = 0; j < n; ++j) {
    long long e0 = 0;
    for (size_t i = 0; i < moduli.size(); ++i) {
      const auto* mp = moduli[i]->ModParams();
      const uint64_t v = cv[i][j].ExportInt(mp);
      const long long c = v > mp->modulus / 2 ? -(long long)(mp->modulus - v) : (long long)v;
      if (i == 0) e0 = c;
      CHECK(c == e0);
    }
    CHECK(e0 % (long long)t == 0);
    const long long e = e0 / (long long)t;
    CHECK(e >= -2 * variance && e <= 2 * variance);
  }
}

// Key generation (rns/rns_relinearization_key_test.cc, rns/rns_galois_key_test.cc):
// the pads are the oracle's SampleUniform from the key's seed, every column
// satisfies the key relation, a Galois key survives Serialize/Deserialize
// (key_a regenerated from the seed on the device), and Compose equals the
// oracle's Galois key switch of the substituted columns.
static void TestKeyGeneration(rlwe::PrngType prng_type) {
  const int log_n = 10, L = 3, n = 1 << log_n, variance = 8;
  const uint64_t t = 65537;
  auto qs = NttPrimes(log_n, L);
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, t).value();
  auto moduli = ctx.MainPrimeModuli();
  OracleChain oc(qs, log_n);
  std::vector<size_t> log_bs(L, 10);
  auto gadget = rlwe::RnsGadget<ModularInt>::Create(
      log_n, log_bs, ctx.MainPrimeModulusComplements(L - 1).value(),
      ctx.MainPrimeModulusCrtFactors(L - 1).value(), moduli).value();
  const int k = gadget.Dimension();
  auto prng = rlwe::SingleThreadChaChaPrng::Create(std::string(32, '\x07')).value();
  ASSIGN_OR_DIE(sk, rlwe::RnsRlweSecretKey<ModularInt>::Sample(log_n, variance, moduli, prng.get()));
  const Poly& s = sk.Key();
  // The secret itself is a centered binomial: s - 0 = 1 * e.
  ASSIGN_OR_DIE(zero, Poly::CreateZero(log_n, moduli));
  CheckKeyColumn(zero, s, zero, zero, moduli, 1, variance);

  ASSIGN_OR_DIE(relin, rlwe::RnsRelinKey<ModularInt>::CreateForBgv(sk, 2, variance, &gadget, t,
                                                                   prng_type));
  CHECK((int)relin.GetKeyA().size() == k && relin.GetPrngType() == prng_type);
  ASSIGN_OR_DIE(s2, s.Mul(s, moduli));
  for (int j = 0; j < k; ++j) {
    Poly target = s2;
    CHECK_OK(target.MulInPlace(gadget.Component(j), moduli));
    CheckKeyColumn(relin.GetKeyA()[j], relin.GetKeyB()[j], s, target, moduli, t, variance);
  }
  if (prng_type == rlwe::PRNG_TYPE_CHACHA) {
    so_chacha_prng* g = (so_chacha_prng*)std::malloc(sizeof(so_chacha_prng));
    CHECK(so_chacha_prng_init(g, (const uint8_t*)relin.PrngSeed().data(), 32) == 0);
    std::vector<uint64_t> a((size_t)L * n);
    for (int j = 0; j < k; ++j) {
      so_sample_uniform_chacha(oc.pp(), L, g, a.data());
      CHECK(Flat(relin.GetKeyA()[j]) == a);
    }
    std::free(g);
  } else {
    auto hp = rlwe::SingleThreadHkdfPrng::Create(relin.PrngSeed()).value();
    for (int j = 0; j < k; ++j) {
      ASSIGN_OR_DIE(a, Poly::SampleUniform(log_n, hp.get(), moduli));
      CHECK(a == relin.GetKeyA()[j]);
    }
  }

  ASSIGN_OR_DIE(g5, rlwe::RnsGaloisKey<ModularInt>::CreateForBgv(sk, 5, variance, &gadget, t,
                                                                 prng_type));
  ASSIGN_OR_DIE(s5, s.Substitute(5, moduli));
  for (int j = 0; j < k; ++j) {
    Poly target = s5;
    CHECK_OK(target.MulInPlace(gadget.Component(j), moduli));
    CheckKeyColumn(g5.GetKeyA()[j], g5.GetKeyB()[j], s, target, moduli, t, variance);
  }
  CHECK(g5.Dimension() == k && g5.SubstitutionPower() == 5);
  ASSIGN_OR_DIE(sg, g5.Serialize());
  rlwe::SerializedRnsGaloisKey sg2;
  CHECK(sg2.ParseFromString(sg.SerializeAsString()));
  ASSIGN_OR_DIE(g5b, rlwe::RnsGaloisKey<ModularInt>::Deserialize(sg2, &gadget, moduli));
  CHECK(g5b.SubstitutionPower() == 5 && g5b.Dimension() == k);
  for (int j = 0; j < k; ++j) {
    CHECK(g5b.GetKeyA()[j] == g5.GetKeyA()[j]);
    CHECK(g5b.GetKeyB()[j] == g5.GetKeyB()[j]);
  }

  // Compose: key for s(X^(5*7)) from keys for s(X^5) and s(X^7).
  ASSIGN_OR_DIE(g7, rlwe::RnsGaloisKey<ModularInt>::CreateForBgv(sk, 7, variance, &gadget, t,
                                                                 prng_type));
  ASSIGN_OR_DIE(g35, g5.Compose(g7));
  CHECK(g35.SubstitutionPower() == 35 && g35.Dimension() == k);
  std::vector<uint64_t> ka((size_t)k * L * n), kb((size_t)k * L * n);
  for (int j = 0; j < k; ++j) {
    auto fa = Flat(g5.GetKeyA()[j]), fb = Flat(g5.GetKeyB()[j]);
    std::copy(fa.begin(), fa.end(), ka.begin() + (size_t)j * L * n);
    std::copy(fb.begin(), fb.end(), kb.begin() + (size_t)j * L * n);
  }
  int lb[64];
  for (int i = 0; i < L; ++i) lb[i] = 10;
  std::vector<uint64_t> sa((size_t)L * n), sb((size_t)L * n), e0((size_t)L * n), e1((size_t)L * n);
  for (int i = 0; i < k; ++i) {
    auto fa = Flat(g7.GetKeyA()[i]), fb = Flat(g7.GetKeyB()[i]);
    so_poly_substitute(oc.pp(), L, 5, fa.data(), sa.data());
    so_poly_substitute(oc.pp(), L, 5, fb.data(), sb.data());
    CHECK(so_galois_apply(oc.pp(), L, lb, sb.data(), sa.data(), 1, ka.data(), kb.data(), e0.data(),
                          e1.data()) == 0);
    CHECK(Flat(g35.GetKeyB()[i]) == e0 && Flat(g35.GetKeyA()[i]) == e1);
  }
  // ComposeWithPads with the digits of sigma(a'_i) equals Compose's b side.
  std::vector<std::vector<Poly>> digits;
  for (int i = 0; i < k; ++i) {
    ASSIGN_OR_DIE(a_sub, g7.GetKeyA()[i].Substitute(5, moduli));
    CHECK_OK(a_sub.ConvertToCoeffForm(moduli));
    ASSIGN_OR_DIE(d, gadget.Decompose(a_sub, moduli));
    for (auto& x : d) CHECK_OK(x.ConvertToNttForm(moduli));
    digits.push_back(std::move(d));
  }
  ASSIGN_OR_DIE(pads, rlwe::RnsGaloisKey<ModularInt>::CreatePadOfComposedKey(g5.GetKeyA(), digits, moduli));
  ASSIGN_OR_DIE(g35p, g5.ComposeWithPads(g7, digits, pads));
  for (int i = 0; i < k; ++i) {
    CHECK(g35p.GetKeyA()[i] == g35.GetKeyA()[i]);
    CHECK(g35p.GetKeyB()[i] == g35.GetKeyB()[i]);
  }
  std::printf("  key generation, Galois serialization and composition (%s): ok\n",
              prng_type == rlwe::PRNG_TYPE_CHACHA ? "ChaCha" : "HKDF");
}

static uint64_t MulMod(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)((unsigned __int128)a * b % m);
}
static uint64_t PowMod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  for (a %= m; e; e >>= 1, a = MulMod(a, a, m))
    if (e & 1) r = MulMod(r, a, m);
  return r;
}

// rns/rns_polynomial.cc:349-852 through the C++ API against the oracle:
// ModReduceMsb, the three basis conversions and both approximate modulus
// reductions (BFV with an auxiliary chain P).
static void TestBfvOps() {
  const int log_n = 11, L = 4, K = 3, n = 1 << log_n;
  auto all = NttPrimes(log_n, L + K);
  std::vector<uint64_t> qs(all.begin(), all.begin() + L), ps(all.begin() + L, all.end());
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, 65537).value();
  auto pctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, ps, 65537).value();
  auto qm = ctx.MainPrimeModuli(), pm = pctx.MainPrimeModuli();
  OracleChain oq(qs, log_n), op(ps, log_n);
  auto imp = [](uint64_t v, const rlwe::PrimeModulus<ModularInt>* m) {
    return ModularInt::ImportInt(v % m->Modulus(), m->ModParams()).value();
  };
  // Constants of the conversion from `from` to `to`.
  struct Consts {
    std::vector<ModularInt> hat_inv, mod_to, inv_from_mod_to_flat;
    std::vector<rlwe::RnsInt<ModularInt>> hat_mod_to, inv_mod_to;
    std::vector<ModularInt> to_mod_from;
  };
  auto consts = [&](const std::vector<uint64_t>& from, const std::vector<const rlwe::PrimeModulus<ModularInt>*>& fm,
                    const std::vector<uint64_t>& to, const std::vector<const rlwe::PrimeModulus<ModularInt>*>& tm) {
    Consts c;
    for (size_t i = 0; i < from.size(); ++i) {
      uint64_t hat = 1;
      for (size_t k = 0; k < from.size(); ++k) if (k != i) hat = MulMod(hat, from[k] % from[i], from[i]);
      c.hat_inv.push_back(imp(PowMod(hat, from[i] - 2, from[i]), fm[i]));
      rlwe::RnsInt<ModularInt> r;
      for (size_t j = 0; j < to.size(); ++j) {
        uint64_t h = 1;
        for (size_t k = 0; k < from.size(); ++k) if (k != i) h = MulMod(h, from[k] % to[j], to[j]);
        r.zs.push_back(imp(h, tm[j]));
      }
      c.hat_mod_to.push_back(r);
      uint64_t t = 1;
      for (uint64_t p : to) t = MulMod(t, p % from[i], from[i]);
      c.to_mod_from.push_back(imp(t, fm[i]));
    }
    for (size_t j = 0; j < to.size(); ++j) {
      uint64_t prod = 1;
      rlwe::RnsInt<ModularInt> r;
      for (size_t i = 0; i < from.size(); ++i) {
        prod = MulMod(prod, from[i] % to[j], to[j]);
        r.zs.push_back(imp(PowMod(from[i] % to[j], to[j] - 2, to[j]), tm[j]));
      }
      c.mod_to.push_back(imp(prod, tm[j]));
      c.inv_mod_to.push_back(r);
    }
    return c;
  };
  auto reps = [](const std::vector<ModularInt>& v) {
    std::vector<uint64_t> r;
    for (const auto& x : v) r.push_back(x.GetMontgomeryRepresentation());
    return r;
  };
  Consts cq = consts(qs, qm, ps, pm), cp = consts(ps, pm, qs, qm);
  std::vector<uint64_t> hat_q_p, inv_q_p;
  for (int i = 0; i < L; ++i) for (int j = 0; j < K; ++j) hat_q_p.push_back(cq.hat_mod_to[i].zs[j].GetMontgomeryRepresentation());
  for (int j = 0; j < K; ++j) for (int i = 0; i < L; ++i) inv_q_p.push_back(cq.inv_mod_to[j].zs[i].GetMontgomeryRepresentation());
  std::vector<uint64_t> hat_p_q;
  for (int j = 0; j < K; ++j) for (int i = 0; i < L; ++i) hat_p_q.push_back(cp.hat_mod_to[j].zs[i].GetMontgomeryRepresentation());

  // ModReduceMsb in both forms.
  auto ql_inv = ctx.MainPrimeModulusInverseResidues()[L - 1].Prefix(L - 1);
  for (bool ntt : {true, false}) {
    auto x = Uniform(oq, L, ntt ? 501 : 502);
    auto p = FromFlat(x, L, n, ntt);
    CHECK_OK(p.ModReduceMsb(ql_inv, qm));
    std::vector<uint64_t> qi;
    for (int i = 0; i < L - 1; ++i) qi.push_back(ql_inv.zs[i].GetMontgomeryRepresentation());
    CHECK(so_poly_mod_reduce_msb(oq.pp(), L, qi.data(), ntt ? 1 : 0, x.data()) == 0);
    x.resize((size_t)(L - 1) * n);
    CHECK(p.NumModuli() == L - 1 && Flat(p) == x);
  }
  // Basis conversions.
  auto x = Uniform(oq, L, 503);
  auto a = FromFlat(x, L, n, false);
  std::vector<uint64_t> e((size_t)K * n);
  ASSIGN_OR_DIE(sw, a.SwitchRnsBasis(qm, pm, cq.hat_inv, cq.hat_mod_to, cq.mod_to));
  so_switch_rns_basis(oq.pp(), L, op.pp(), K, reps(cq.hat_inv).data(), hat_q_p.data(),
                      reps(cq.mod_to).data(), x.data(), e.data());
  CHECK(Flat(sw) == e && !sw.IsNttForm());
  ASSIGN_OR_DIE(sc, a.ScaleAndSwitchRnsBasis(qm, pm, cq.hat_inv, cq.inv_mod_to, cq.to_mod_from));
  so_scale_and_switch_rns_basis(oq.pp(), L, op.pp(), K, reps(cq.hat_inv).data(), inv_q_p.data(),
                                reps(cq.to_mod_from).data(), x.data(), e.data());
  CHECK(Flat(sc) == e);
  auto an = FromFlat(x, L, n, true);
  for (bool bal : {false, true}) {
    ASSIGN_OR_DIE(ap, an.ApproxSwitchRnsBasis(qm, pm, cq.hat_inv, cq.hat_mod_to, bal));
    so_approx_switch_rns_basis(oq.pp(), L, op.pp(), K, reps(cq.hat_inv).data(), hat_q_p.data(), 1,
                               bal ? 1 : 0, x.data(), e.data());
    CHECK(Flat(ap) == e);
  }
  CHECK(!an.SwitchRnsBasis(qm, pm, cq.hat_inv, cq.hat_mod_to, cq.mod_to).ok());  // NTT form
  // Approximate modulus reductions by P (aux chain), a in NTT form mod Q.
  const uint64_t t = 65537;
  uint64_t P_mod_t = 1;
  for (uint64_t p : ps) P_mod_t = MulMod(P_mod_t, p % t, t);
  std::vector<ModularInt> p_inv, t_inv;
  for (int i = 0; i < L; ++i) {
    uint64_t P = 1;
    for (uint64_t p : ps) P = MulMod(P, p % qs[i], qs[i]);
    p_inv.push_back(imp(PowMod(P, qs[i] - 2, qs[i]), qm[i]));
  }
  for (int j = 0; j < K; ++j) t_inv.push_back(imp(PowMod(t, ps[j] - 2, ps[j]), pm[j]));
  for (int lsb = 0; lsb < 2; ++lsb) {
    auto av = Uniform(oq, L, 600 + lsb), auxv = Uniform(op, K, 610 + lsb);
    auto pa = FromFlat(av, L, n, true);
    auto paux = FromFlat(auxv, K, n, false);
    std::vector<uint64_t> delta((size_t)L * n), aux2 = auxv;
    if (lsb) {
      CHECK_OK(pa.ApproxModReduceLsb(paux, qm, pm, p_inv, cp.hat_inv, cp.hat_mod_to, t_inv, P_mod_t, t));
      std::vector<uint64_t> sc_aux;
      for (int j = 0; j < K; ++j) {
        const so_mont* mj = &op.primes[j]->mp;
        sc_aux.push_back(so_mul(mj, so_import(mj, P_mod_t), so_neg(mj, t_inv[j].GetMontgomeryRepresentation())));
      }
      so_poly_mul_scalar(op.pp(), K, aux2.data(), sc_aux.data());
      so_approx_switch_rns_basis(op.pp(), K, oq.pp(), L, reps(cp.hat_inv).data(), hat_p_q.data(), 0, 1,
                                 aux2.data(), delta.data());
      so_poly_ntt_fwd(oq.pp(), L, delta.data());
      so_poly_mul_int(oq.pp(), L, delta.data(), t);
      so_poly_mul_int(oq.pp(), L, av.data(), P_mod_t);
      so_poly_pointwise(oq.pp(), L, SO_ADD, av.data(), delta.data(), nullptr);
    } else {
      CHECK_OK(pa.ApproxModReduceMsb(paux, qm, pm, p_inv, cp.hat_inv, cp.hat_mod_to));
      so_approx_switch_rns_basis(op.pp(), K, oq.pp(), L, reps(cp.hat_inv).data(), hat_p_q.data(), 0, 0,
                                 aux2.data(), delta.data());
      so_poly_ntt_fwd(oq.pp(), L, delta.data());
      so_poly_pointwise(oq.pp(), L, SO_SUB, av.data(), delta.data(), nullptr);
    }
    so_poly_mul_scalar(oq.pp(), L, av.data(), reps(p_inv).data());
    CHECK(Flat(pa) == av);
  }
  std::printf("  BFV: ModReduceMsb, basis conversions, approximate modulus reductions: ok\n");
}

// rns/rns_ciphertext.h:309-495 (FusedAbsorb*WithoutPad*, MergeLazyOperations,
// SetPadComponent) over rns/lazy_rns_polynomial.cc:89-254: lazy sums,
// differences and products of "b" parts against the oracle's 128-bit lazy
// sum, the eager absorb against the lazy one, and the error bookkeeping.
static void TestLazyAbsorb(int log_n, int L) {
  using Rlwe = rlwe::RnsRlweCiphertext<ModularInt>;
  const int n = 1 << log_n;
  auto qs = NttPrimes(log_n, L);
  auto ctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, qs, 65537).value();
  auto moduli = ctx.MainPrimeModuli();
  OracleChain oc(qs, log_n);
  auto error_params = rlwe::RnsErrorParams<ModularInt>::Create(log_n, moduli, {}, 17, 4.0).value();
  const double bp = error_params.B_plaintext();
  CHECK(bp == 17.0 * 17.0 * std::sqrt(3.0 * n));
  auto make = [&](uint64_t seed, double err, int len) {
    std::vector<Poly> comps;
    for (int i = 0; i < len; ++i) comps.push_back(FromFlat(Uniform(oc, L, seed + i), L, n, true));
    return Rlwe(std::move(comps), moduli, 1, err, &error_params);
  };
  const int T = 7;
  const size_t words = (size_t)L * n;
  std::vector<Rlwe> c1s, c2s;
  std::vector<Poly> ptxts;
  std::vector<uint64_t> as(T * words), bs(T * words), cs(T * words);
  std::vector<int> kinds(T);
  for (int t = 0; t < T; ++t) {
    c1s.push_back(make(3000 + 10 * t, 1.0 + t, 2));
    c2s.push_back(make(3005 + 10 * t, 0.5 + t, 2));
    auto pt = Uniform(oc, L, 4000 + t);
    ptxts.push_back(FromFlat(pt, L, n, true));
    auto b1 = Flat(c1s[t].Components()[0]), b2 = Flat(c2s[t].Components()[0]);
    kinds[t] = t % 3;
    std::copy(b1.begin(), b1.end(), as.begin() + t * words);
    std::copy((kinds[t] ? b2 : pt).begin(), (kinds[t] ? b2 : pt).end(), bs.begin() + t * words);
    std::copy(pt.begin(), pt.end(), cs.begin() + t * words);
  }
  Rlwe acc = make(5000, 2.0, 2);
  const auto b0 = Flat(acc.Components()[0]), a0 = Flat(acc.Components()[1]);
  auto st = acc.MergeLazyOperations();
  CHECK(!st.ok() && st.message() == "The ciphertext was not in lazy mode.");
  double err = 2.0;
  for (int t = 0; t < T; ++t) {
    if (kinds[t] == 0) {
      CHECK_OK(acc.FusedAbsorbAddInPlaceWithoutPadLazily(c1s[t], ptxts[t]));
      err += c1s[t].Error() * bp;
    } else if (kinds[t] == 1) {
      CHECK_OK(acc.FusedAbsorbSumAddInPlaceWithoutPadLazily(c1s[t], c2s[t], ptxts[t]));
      err += (c1s[t].Error() + c2s[t].Error()) * bp;
    } else {
      CHECK_OK(acc.FusedAbsorbDiffAddInPlaceWithoutPadLazily(c1s[t], c2s[t], ptxts[t]));
      err += (c1s[t].Error() + c2s[t].Error()) * bp;
    }
  }
  CHECK(Flat(acc.Components()[0]) == b0);  // the lazy part is not merged yet
  CHECK_OK(acc.MergeLazyOperations());
  std::vector<uint64_t> lazy(words);
  so_lazy_mixed(oc.pp(), L, T, kinds.data(), as.data(), bs.data(), cs.data(), lazy.data());
  std::vector<uint64_t> expect(words);
  for (size_t j = 0; j < words; ++j) expect[j] = so_add(&oc.primes[j / n]->mp, b0[j], lazy[j]);
  CHECK(Flat(acc.Components()[0]) == expect);
  CHECK(Flat(acc.Components()[1]) == a0);
  CHECK(acc.Error() == err);
  CHECK(!acc.MergeLazyOperations().ok());

  // A difference first: the call that creates the lazy part leaves the error.
  Rlwe d = make(5100, 3.0, 2);
  CHECK_OK(d.FusedAbsorbDiffAddInPlaceWithoutPadLazily(c1s[0], c2s[0], ptxts[0]));
  CHECK(d.Error() == 3.0);

  // Eager and lazy products agree.
  Rlwe eager = make(6000, 1.0, 2), lz = make(6000, 1.0, 2);
  for (int t = 0; t < 3; ++t) {
    CHECK_OK(eager.FusedAbsorbAddInPlaceWithoutPad(c1s[t], ptxts[t]));
    CHECK_OK(lz.FusedAbsorbAddInPlaceWithoutPadLazily(c1s[t], ptxts[t]));
  }
  CHECK_OK(lz.MergeLazyOperations());
  CHECK(eager.Components()[0] == lz.Components()[0] && eager.Components()[1] == lz.Components()[1]);
  CHECK(eager.Error() == lz.Error());

  // Lazy absorbs into a ciphertext with no components, then the pad.
  Rlwe z(std::vector<Poly>{}, moduli, 1, 0.0, &error_params);
  CHECK_OK(z.FusedAbsorbAddInPlaceWithoutPadLazily(c1s[0], ptxts[0]));
  CHECK_OK(z.MergeLazyOperations());
  CHECK(z.Degree() == 0);
  std::vector<uint64_t> e0(words);
  so_lazy_fma(oc.pp(), L, 1, as.data(), bs.data(), e0.data());
  CHECK(Flat(z.Components()[0]) == e0);
  CHECK_OK(z.SetPadComponent(c2s[0].Components()[1]));
  CHECK(z.Degree() == 1 && z.Components()[1] == c2s[0].Components()[1]);
  CHECK_OK(z.SetPadComponent(c2s[1].Components()[1]));
  CHECK(z.Degree() == 1 && z.Components()[1] == c2s[1].Components()[1]);

  // Reference error messages.
  Rlwe deg2 = make(7000, 1.0, 3);
  st = deg2.FusedAbsorbAddInPlaceWithoutPad(c1s[0], ptxts[0]);
  CHECK(!st.ok() && st.message() == "`FusedAbsorbAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
  st = acc.FusedAbsorbAddInPlaceWithoutPadLazily(deg2, ptxts[0]);
  CHECK(!st.ok() && st.message() == "`FusedAbsorbAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
  st = acc.FusedAbsorbSumAddInPlaceWithoutPadLazily(c1s[0], deg2, ptxts[0]);
  CHECK(!st.ok() && st.message() == "`FusedAbsorbSumAddInPlaceWithoutPad only applies to degree 1 ciphertext.");
  CHECK(!deg2.SetPadComponent(ptxts[0]).ok());
  Rlwe other_power(std::vector<Poly>{c1s[0].Components()[0], c1s[0].Components()[1]}, moduli, 3,
                   1.0, &error_params);
  st = acc.FusedAbsorbAddInPlaceWithoutPad(other_power, ptxts[0]);
  CHECK(!st.ok() && st.message() == "`ctxt` has a mismatched key power.");
  auto coeff = ptxts[0];
  CHECK_OK(coeff.ConvertToCoeffForm(moduli));
  auto lz_bad = rlwe::LazyRnsPolynomial<ModularInt>::Create(coeff, ptxts[1], moduli);
  CHECK(!lz_bad.ok() && lz_bad.status().message() == "Polynomials `a` and `b` must be in the NTT form.");
  auto lzero = rlwe::LazyRnsPolynomial<ModularInt>::CreateZero(log_n, moduli).value();
  CHECK(Flat(lzero.Export(moduli).value()) == std::vector<uint64_t>(words, 0));
  CHECK(!lzero.Export(rlwe::ModuliSpan<ModularInt>(moduli.data(), L - 1)).ok());

  // Moduli of 2^60 and above: no fused multiply-sum-add.
  auto wide = NttPrimes(log_n, 1, 62);
  auto wctx = rlwe::RnsContext<ModularInt>::CreateForBgv(log_n, wide, 65537).value();
  auto wm = wctx.MainPrimeModuli();
  OracleChain woc(wide, log_n);
  auto wp = FromFlat(Uniform(woc, 1, 9), 1, n, true);
  auto wl = rlwe::LazyRnsPolynomial<ModularInt>::Create(wp, wp, wm).value();
  st = wl.FusedMulSumAddInPlace(wp, wp, wp, wm);
  CHECK(!st.ok() && st.message() == "Modulus is too large to perform fused multiply-sum-add.");
  CHECK_OK(wl.FusedMulAddInPlace(wp, wp, wm));
}

int main() {
  TestMontgomery();
  TestNttParameters();
  TestRnsPolynomial();
  TestModReduceAndGadget();
  TestKeySwitching(10, 3, 4);
  TestKeySwitching(14, 8, 2);
  TestSerialization();
  TestKeyGeneration(rlwe::PRNG_TYPE_CHACHA);
  TestKeyGeneration(rlwe::PRNG_TYPE_HKDF);
  TestBfvOps();
  TestLazyAbsorb(10, 3);
  TestLazyAbsorb(14, 8);
  std::printf("rns_api_test: %d checks passed\n", g_checks);
  return 0;
}
