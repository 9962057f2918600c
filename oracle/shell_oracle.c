/*
 * shell_oracle.c -- CPU restatement of SHELL's RNS polynomial hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the HIP path is tested against it; it is never
 * part of the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it (through oracle/oracle.py).
 *
 * Each routine restates the reference routine cited beside it, step for step,
 * including lazy reduction, so that the CPU baseline timed from it runs the
 * reference's own algorithm (scalar, single thread).  The reference cannot be
 * compiled here (it needs Abseil, Highway, protobuf and Bazel), so this file
 * is written from the reference sources read as text.
 */
#include "shell_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* L1: MontgomeryIntParams<Uint64> and MontgomeryInt<Uint64>                */
/* ------------------------------------------------------------------------ */

static unsigned bit_length(uint64_t x) {  /* bits_util.h:102-104 */
  unsigned b = 0;
  while (x) { b++; x >>= 1; }
  return b;
}

/* MontgomeryIntParams::Create + ctor + Inverses
 * (montgomery.cc:44-59, montgomery.h:177-192, montgomery.cc:62-92). */
int so_mont_init(uint64_t q, so_mont* p) {
  if ((q >> 62) != 0) return -1;  /* "The modulus should be less than 2^62." */
  if ((q % 2) == 0) return -2;    /* "The modulus should be odd." */
  memset(p, 0, sizeof(*p));
  const u128 r = (u128)1 << 64;
  p->modulus = q;
  p->log_modulus = bit_length(q);
  p->r_mod_modulus = (uint64_t)(r % q);
  p->r_mod_modulus_barrett = (uint64_t)(((u128)p->r_mod_modulus << 64) / q);
  /* Inverses: invariant x*2^w - y*q is a power of two, from 2^w down to 1. */
  u128 x = 1, y = 0;
  for (int i = 64; i > 0; i--) {
    if (x & 1) { x += q; y += r; }
    x >>= 1;
    y >>= 1;
  }
  p->inv_r = (uint64_t)x;
  p->inv_modulus = (uint64_t)y;
  p->inv_r_barrett = (uint64_t)(((u128)p->inv_r << 64) / q);
  p->barrett_numerator = (uint64_t)(r / q);
  u128 big = ((u128)1 << 127) / q;
  p->barrett_numerator_big_lo = (uint64_t)big;
  p->barrett_numerator_big_hi = (uint64_t)(big >> 64);
  return 0;
}

/* MontgomeryIntParams::BarrettReduce (montgomery.h:134-141). */
uint64_t so_barrett_reduce(const so_mont* p, uint64_t in) {
  uint64_t out = (uint64_t)(((u128)p->barrett_numerator * in) >> 64);
  out = in - out * p->modulus;
  return (out >= p->modulus) ? out - p->modulus : out;
}

/* (a * b) >> 127 for 128-bit a, b, via four 64x64 partial products. */
static u128 mul_shr127(u128 a, u128 b) {
  uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64);
  uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
  u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0,
       p11 = (u128)a1 * b1;
  u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;  /* < 3*2^64 */
  u128 hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);  /* bits 128.. */
  uint64_t w1 = (uint64_t)mid;  /* bits 64..127 */
  /* result = floor(product / 2^127) = hi*2 + (w1 >> 63) */
  return (hi << 1) | (w1 >> 63);
}

/* MontgomeryIntParams::BarrettReduceBigInt (montgomery.h:143-152). */
uint64_t so_barrett_reduce_big(const so_mont* p, uint64_t lo, uint64_t hi) {
  u128 in = ((u128)hi << 64) | lo;
  u128 bn = ((u128)p->barrett_numerator_big_hi << 64) |
            p->barrett_numerator_big_lo;
  uint64_t out = (uint64_t)mul_shr127(bn, in);
  out = (uint64_t)in - out * p->modulus;
  return (out >= p->modulus) ? out - p->modulus : out;
}

/* MontgomeryIntParams::ExportInt (montgomery.h:156-164). */
uint64_t so_export(const so_mont* p, uint64_t in) {
  uint64_t out = (uint64_t)(((u128)p->inv_r_barrett * in) >> 64);
  out = in * p->inv_r - out * p->modulus;
  return (out >= p->modulus) ? out - p->modulus : out;
}

/* ToMontgomeryInt / ImportInt (montgomery.cc:26-40, 94-98). */
uint64_t so_import(const so_mont* p, uint64_t n) {
  uint64_t result = (uint64_t)(((u128)p->r_mod_modulus_barrett * n) >> 64);
  result = n * p->r_mod_modulus - result * p->modulus;
  result -= (result >= p->modulus) ? p->modulus : 0;
  return result;
}

/* MontgomeryInt::MulInPlace, REDC (montgomery.h:318-336). */
uint64_t so_mul(const so_mont* p, uint64_t a, uint64_t b) {
  u128 product = (u128)a * b;
  uint64_t u = (uint64_t)product * p->inv_modulus;
  u128 t = product + (u128)p->modulus * u;
  uint64_t t_msb = (uint64_t)(t >> 64);
  return (t_msb >= p->modulus) ? t_msb - p->modulus : t_msb;
}

/* MontgomeryInt::MulConstantInPlace, Shoup (montgomery.h:362-372). */
uint64_t so_mul_const(const so_mont* p, uint64_t x, uint64_t c, uint64_t cb) {
  uint64_t out = (uint64_t)(((u128)cb * x) >> 64);
  uint64_t n = x * c - out * p->modulus;
  n -= (n >= p->modulus) ? p->modulus : 0;
  return n;
}

/* AddInPlace / SubInPlace / NegateInPlace (montgomery.h:383-415). */
uint64_t so_add(const so_mont* p, uint64_t a, uint64_t b) {
  return so_barrett_reduce(p, a + b);
}
uint64_t so_sub(const so_mont* p, uint64_t a, uint64_t b) {
  return so_barrett_reduce(p, a + (p->modulus - b));
}
uint64_t so_neg(const so_mont* p, uint64_t a) { return p->modulus - a; }

/* FusedMulAddInPlace (montgomery.h:446-465). */
uint64_t so_fma(const so_mont* p, uint64_t acc, uint64_t a, uint64_t b) {
  u128 t = (u128)p->r_mod_modulus * acc + (u128)a * b;
  uint64_t u = (uint64_t)t * p->inv_modulus;
  t += (u128)p->modulus * u;
  uint64_t t_msb = (uint64_t)(t >> 64);
  return (t_msb >= p->modulus) ? t_msb - p->modulus : t_msb;
}

/* FusedMulConstantAddInPlace (montgomery.h:467-506). */
uint64_t so_fma_const(const so_mont* p, uint64_t acc, uint64_t a, uint64_t c,
                      uint64_t cb) {
  uint64_t out = (uint64_t)(((u128)cb * a + (u128)p->barrett_numerator * acc) >>
                            64);
  acc += a * c - out * p->modulus;
  acc -= (acc >= p->modulus) ? p->modulus : 0;
  return acc;
}

/* MontgomeryInt::GetConstant (montgomery.cc:221-227). */
void so_get_constant(const so_mont* p, uint64_t w, uint64_t* c, uint64_t* cb) {
  uint64_t constant = so_export(p, w);
  *c = constant;
  *cb = (uint64_t)(((u128)constant << 64) / p->modulus);
}

/* MontgomeryInt::ModExp (montgomery.cc:473-493). */
uint64_t so_modexp(const so_mont* p, uint64_t base, uint64_t e) {
  uint64_t result = p->r_mod_modulus;  /* ImportOne */
  while (e > 0) {
    if (e % 2 == 1) result = so_mul(p, result, base);
    base = so_mul(p, base, base);
    e >>= 1;
  }
  return result;
}

/* MultiplicativeInverse (montgomery.cc:495-499). */
uint64_t so_inverse(const so_mont* p, uint64_t a) {
  return so_modexp(p, a, p->modulus - 2);
}

/* ------------------------------------------------------------------------ */
/* L2: NTT parameters (ntt_parameters.h, ntt_parameters.cc)                 */
/* ------------------------------------------------------------------------ */

/* Bitrev (ntt_parameters.cc:21-31). */
static unsigned bitrev(unsigned in, unsigned log_n) {
  unsigned out = 0;
  for (unsigned i = 0; i < log_n; i++) {
    out <<= 1;
    out |= in & 1;
    in >>= 1;
  }
  return out;
}

/* PrimitiveNthRootOfUnity (ntt_parameters.h:41-71). */
static int primitive_root(const so_mont* p, unsigned log_n, uint64_t* root) {
  uint64_t n = (uint64_t)1 << log_n;
  uint64_t half_n = n >> 1;
  uint64_t k = (p->modulus - 1) / n;
  uint64_t one = p->r_mod_modulus;
  for (uint64_t t = 2; t < p->modulus; t++) {
    uint64_t mt = so_import(p, t);
    uint64_t cand = so_modexp(p, mt, k);
    if (so_modexp(p, cand, half_n) != one) {
      *root = cand;
      return 0;
    }
  }
  return -1;
}

/* InitializeNttParameters (ntt_parameters.h:183-235) with NttPsis,
 * NttPsisBitrev, NttPsisInvBitrev (ntt_parameters.h:79-157). */
so_prime* so_prime_create(uint64_t q, int log_n, int* status) {
  *status = 0;
  if (log_n <= 0) { *status = -10; return NULL; }   /* log_n must be positive */
  if (log_n > 15) { *status = -11; return NULL; }   /* kMaxLogNumCoeffs */
  so_prime* pr = (so_prime*)calloc(1, sizeof(so_prime));
  if (so_mont_init(q, &pr->mp) != 0) { *status = -1; free(pr); return NULL; }
  const so_mont* p = &pr->mp;
  unsigned n = 1u << log_n;
  uint64_t two_n = (uint64_t)1 << (log_n + 1);
  if (q % two_n != 1) { *status = -12; free(pr); return NULL; }
  pr->log_n = log_n;
  pr->n_inv = so_inverse(p, so_import(p, n));
  if (primitive_root(p, (unsigned)log_n + 1, &pr->psi) != 0) {
    *status = -13; free(pr); return NULL;
  }
  uint64_t* row = (uint64_t*)malloc(sizeof(uint64_t) * n);
  for (unsigned i = 0; i < n; i++) row[i] = so_add(p, 0, so_modexp(p, pr->psi, i));
  pr->bitrevs = (uint32_t*)malloc(sizeof(uint32_t) * n);
  for (unsigned i = 0; i < n; i++) pr->bitrevs[i] = bitrev(i, (unsigned)log_n);
  pr->psis_bitrev = (uint64_t*)malloc(sizeof(uint64_t) * n);
  for (unsigned i = 0; i < n; i++) pr->psis_bitrev[i] = row[pr->bitrevs[i]];
  /* inverse table: reverse entries 1..n-1, bit-reverse, then scale. */
  uint64_t* rev = (uint64_t*)malloc(sizeof(uint64_t) * n);
  rev[0] = row[0];
  for (unsigned i = 1; i < n; i++) rev[i] = row[n - i];
  uint64_t psi_inv = so_neg(p, rev[1 % n]);
  uint64_t negative_psi_inv = rev[1 % n];
  pr->psis_inv_bitrev = (uint64_t*)malloc(sizeof(uint64_t) * n);
  for (unsigned i = 0; i < n; i++) pr->psis_inv_bitrev[i] = rev[pr->bitrevs[i]];
  pr->psis_inv_bitrev[0] = so_mul(p, pr->psis_inv_bitrev[0], psi_inv);
  for (unsigned i = 1; i < n; i++)
    pr->psis_inv_bitrev[i] = so_mul(p, pr->psis_inv_bitrev[i], negative_psi_inv);
  free(rev);
  free(row);
  pr->psis_bitrev_c = (uint64_t*)malloc(sizeof(uint64_t) * n);
  pr->psis_bitrev_cb = (uint64_t*)malloc(sizeof(uint64_t) * n);
  pr->psis_inv_bitrev_c = (uint64_t*)malloc(sizeof(uint64_t) * n);
  pr->psis_inv_bitrev_cb = (uint64_t*)malloc(sizeof(uint64_t) * n);
  for (unsigned i = 0; i < n; i++) {
    so_get_constant(p, pr->psis_bitrev[i], &pr->psis_bitrev_c[i],
                    &pr->psis_bitrev_cb[i]);
    so_get_constant(p, pr->psis_inv_bitrev[i], &pr->psis_inv_bitrev_c[i],
                    &pr->psis_inv_bitrev_cb[i]);
  }
  return pr;
}

void so_prime_destroy(so_prime* pr) {
  if (!pr) return;
  free(pr->psis_bitrev); free(pr->psis_bitrev_c); free(pr->psis_bitrev_cb);
  free(pr->psis_inv_bitrev); free(pr->psis_inv_bitrev_c);
  free(pr->psis_inv_bitrev_cb); free(pr->bitrevs);
  free(pr);
}

const so_mont* so_prime_mont(const so_prime* pr) { return &pr->mp; }

/* Copies the six n-word tables (psis_bitrev, its (c, cb) constants,
 * psis_inv_bitrev, its (c, cb) constants), the bit-reverse table and
 * (psi, n_inv). */
void so_prime_tables(const so_prime* pr, uint64_t* out6n, uint32_t* bitrevs,
                     uint64_t* psi_ninv) {
  size_t n = (size_t)1 << pr->log_n;
  memcpy(out6n + 0 * n, pr->psis_bitrev, n * 8);
  memcpy(out6n + 1 * n, pr->psis_bitrev_c, n * 8);
  memcpy(out6n + 2 * n, pr->psis_bitrev_cb, n * 8);
  memcpy(out6n + 3 * n, pr->psis_inv_bitrev, n * 8);
  memcpy(out6n + 4 * n, pr->psis_inv_bitrev_c, n * 8);
  memcpy(out6n + 5 * n, pr->psis_inv_bitrev_cb, n * 8);
  memcpy(bitrevs, pr->bitrevs, n * 4);
  psi_ninv[0] = pr->psi;
  psi_ninv[1] = pr->n_inv;
}

/* ------------------------------------------------------------------------ */
/* L2: transforms (dft_transformations.h)                                   */
/* ------------------------------------------------------------------------ */

/* IterativeCooleyTukey (dft_transformations.h:59-100): natural order in,
 * bit-reversed order out; odd layers lazy, even layers Barrett-reduced. */
void so_ntt_fwd(const so_prime* pr, uint64_t* x) {
  const so_mont* p = &pr->mp;
  const uint64_t q = p->modulus;
  const int n = 1 << pr->log_n;
  int index_psi = 1;
  for (int i = pr->log_n - 1; i >= 0; i--) {
    const int half_m = 1 << i;
    const int m = half_m << 1;
    for (int k = 0; k < n; k += m) {
      const uint64_t c = pr->psis_bitrev_c[index_psi];
      const uint64_t cb = pr->psis_bitrev_cb[index_psi];
      for (int j = 0; j < half_m; j++) {
        uint64_t t = so_mul_const(p, x[k + j + half_m], c, cb);
        uint64_t u = x[k + j];
        if (i & 1) {
          x[k + j] = u + t;             /* LazyAddInPlace */
          x[k + j + half_m] = u + (q - t);  /* LazySubInPlace */
        } else {
          x[k + j] = so_add(p, u, t);
          x[k + j + half_m] = so_sub(p, u, t);
        }
      }
      index_psi++;
    }
  }
}

/* IterativeGentlemanSande + normalization by n^-1
 * (dft_transformations.h:118-153, 181-201). */
void so_ntt_inv(const so_prime* pr, uint64_t* x) {
  const so_mont* p = &pr->mp;
  const uint64_t q = p->modulus;
  const int n = 1 << pr->log_n;
  int index_psi_inv = 0;
  for (int i = 0; i < pr->log_n; i++) {
    const int half_m = 1 << i;
    const int m = half_m << 1;
    for (int k = 0; k < n; k += m) {
      const uint64_t c = pr->psis_inv_bitrev_c[index_psi_inv];
      const uint64_t cb = pr->psis_inv_bitrev_cb[index_psi_inv];
      for (int j = 0; j < half_m; j++) {
        uint64_t t = x[k + j + half_m];
        uint64_t u = x[k + j];
        x[k + j] = so_add(p, u, t);
        x[k + j + half_m] = so_mul_const(p, u + (q - t), c, cb);
      }
      index_psi_inv++;
    }
  }
  for (int j = 0; j < n; j++) x[j] = so_mul(p, x[j], pr->n_inv);
}

/* ------------------------------------------------------------------------ */
/* L3: RnsPolynomial operations (rns/rns_polynomial.{h,cc})                 */
/* ------------------------------------------------------------------------ */

/* NegateInPlace / AddInPlace / SubInPlace / MulInPlace(poly) /
 * FusedMulAddInPlace (rns/rns_polynomial.h:159-365) limb by limb. */
void so_poly_pointwise(so_prime* const* primes, int nl, int op, uint64_t* x,
                       const uint64_t* a, const uint64_t* b) {
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    const size_t n = (size_t)1 << primes[i]->log_n;
    uint64_t* xi = x + i * n;
    const uint64_t* ai = a ? a + i * n : NULL;
    const uint64_t* bi = b ? b + i * n : NULL;
    for (size_t j = 0; j < n; j++) {
      switch (op) {
        case SO_ADD: xi[j] = so_add(p, xi[j], ai[j]); break;
        case SO_SUB: xi[j] = so_sub(p, xi[j], ai[j]); break;
        case SO_MUL: xi[j] = so_mul(p, xi[j], ai[j]); break;
        case SO_NEG: xi[j] = so_neg(p, xi[j]); break;
        case SO_FMA: xi[j] = so_fma(p, xi[j], ai[j], bi[j]); break;
        default: break;
      }
    }
  }
}

/* MulInPlace(Span<const ModularInt>) (rns/rns_polynomial.h:265-283). */
void so_poly_mul_scalar(so_prime* const* primes, int nl, uint64_t* x,
                        const uint64_t* scalars) {
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    const size_t n = (size_t)1 << primes[i]->log_n;
    for (size_t j = 0; j < n; j++) x[i * n + j] = so_mul(p, x[i * n + j], scalars[i]);
  }
}

/* MulInPlace(Int scalar): ImportInt per limb then BatchMulInPlace
 * (rns/rns_polynomial.h:243-258). */
void so_poly_mul_int(so_prime* const* primes, int nl, uint64_t* x,
                     uint64_t scalar) {
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    const size_t n = (size_t)1 << primes[i]->log_n;
    const uint64_t s = so_import(p, scalar);
    for (size_t j = 0; j < n; j++) x[i * n + j] = so_mul(p, x[i * n + j], s);
  }
}

/* ConvertToNttForm / ConvertToCoeffForm (rns/rns_polynomial.cc:171-207). */
void so_poly_ntt_fwd(so_prime* const* primes, int nl, uint64_t* x) {
  for (int i = 0; i < nl; i++)
    so_ntt_fwd(primes[i], x + ((size_t)i << primes[i]->log_n));
}
void so_poly_ntt_inv(so_prime* const* primes, int nl, uint64_t* x) {
  for (int i = 0; i < nl; i++)
    so_ntt_inv(primes[i], x + ((size_t)i << primes[i]->log_n));
}

/* RnsPolynomial::Substitute (rns/rns_polynomial.cc:209-260).  The running
 * index psi_power_index carries over from one limb to the next, exactly as
 * in the reference loop. */
int so_poly_substitute(so_prime* const* primes, int nl, int power,
                       const uint64_t* in, uint64_t* out) {
  const int num_coeffs = 1 << primes[0]->log_n;
  if (power < 0 || (power % 2) == 0 || power >= 2 * num_coeffs) return -1;
  int psi_power_index = (power - 1) / 2;
  memcpy(out, in, sizeof(uint64_t) * (size_t)nl * num_coeffs);
  for (int i = 0; i < nl; i++) {
    const uint32_t* brv = primes[i]->bitrevs;
    const uint64_t* src = in + (size_t)i * num_coeffs;
    uint64_t* dst = out + (size_t)i * num_coeffs;
    for (int j = 0; j < num_coeffs; j++) {
      int coeff_index = (int)brv[psi_power_index];
      if (coeff_index >= num_coeffs) return -2;
      dst[brv[j]] = src[coeff_index];
      psi_power_index = (psi_power_index + power) % num_coeffs;
    }
  }
  return 0;
}

/* ConvertModulusBalanced (modulus_conversion.h:30-83). */
void so_convert_modulus_balanced(const so_mont* pq, const so_mont* pp,
                                 const uint64_t* in, uint64_t* out, int n) {
  const uint64_t q = pq->modulus, pmod = pp->modulus;
  if (q == pmod) {
    memmove(out, in, sizeof(uint64_t) * (size_t)n);
    return;
  }
  const uint64_t q_by_2 = q >> 1;
  const uint64_t diff = q > pmod ? q - pmod : pmod - q;
  const uint64_t diff_p = so_import(pp, diff);
  for (int i = 0; i < n; i++) {
    uint64_t x = so_export(pq, in[i]);
    uint64_t xp = so_import(pp, x);
    if (x > q_by_2) xp = (pmod > q) ? so_add(pp, xp, diff_p) : so_sub(pp, xp, diff_p);
    out[i] = xp;
  }
}

/* RnsPolynomial::ModReduceLsb (rns/rns_polynomial.cc:262-344).  On return
 * the first (nl - 1) limbs of x hold the result. */
int so_poly_mod_reduce_lsb(so_prime* const* primes, int nl, uint64_t t,
                           const uint64_t* ql_inv, int is_ntt, uint64_t* x) {
  if (nl <= 1) return -1;
  const size_t n = (size_t)1 << primes[0]->log_n;
  const int nr = nl - 1;
  const so_prime* pl = primes[nl - 1];
  const so_mont* ml = &pl->mp;
  uint64_t* delta_ql = (uint64_t*)malloc(n * 8);
  memcpy(delta_ql, x + (size_t)nr * n, n * 8);  /* DetachLastCoeffVector */
  uint64_t r = ml->modulus % t;
  uint64_t k = (ml->modulus - r) / t;
  uint64_t k_mod_ql = so_import(ml, k);
  for (size_t j = 0; j < n; j++) delta_ql[j] = so_mul(ml, delta_ql[j], k_mod_ql);
  so_poly_mul_int(primes, nr, x, r);
  uint64_t* delta = (uint64_t*)malloc(n * 8 * (size_t)nr);
  if (is_ntt) {
    so_ntt_inv(pl, delta_ql);
    for (int i = 0; i < nr; i++)
      so_convert_modulus_balanced(ml, &primes[i]->mp, delta_ql, delta + i * n, (int)n);
    so_poly_ntt_fwd(primes, nr, delta);
    so_poly_mul_int(primes, nr, delta, t);
    so_poly_pointwise(primes, nr, SO_ADD, x, delta, NULL);
    so_poly_mul_scalar(primes, nr, x, ql_inv);
  } else {
    for (int i = 0; i < nr; i++) {
      const so_mont* mi = &primes[i]->mp;
      uint64_t* di = delta + i * n;
      so_convert_modulus_balanced(ml, mi, delta_ql, di, (int)n);
      uint64_t t_mod_qi = so_import(mi, t);
      for (size_t j = 0; j < n; j++) di[j] = so_mul(mi, di[j], t_mod_qi);
      for (size_t j = 0; j < n; j++) x[i * n + j] = so_add(mi, x[i * n + j], di[j]);
      for (size_t j = 0; j < n; j++) x[i * n + j] = so_mul(mi, x[i * n + j], ql_inv[i]);
    }
  }
  free(delta);
  free(delta_ql);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Gadget (gadget.h, rns/rns_gadget.cc)                                     */
/* ------------------------------------------------------------------------ */

/* GadgetSize (gadget.h:43-47). */
int so_gadget_dims(so_prime* const* primes, int nl, const int* log_bs,
                   int* dims) {
  int k = 0;
  for (int i = 0; i < nl; i++) {
    dims[i] = ((int)primes[i]->mp.log_modulus + (log_bs[i] - 1)) / log_bs[i];
    k += dims[i];
  }
  return k;
}

/* RnsGadget::Create (rns/rns_gadget.cc:42-102): gs is k x nl words, entry
 * (j, limb) is the Montgomery rep of component j modulo q_limb. */
void so_gadget_components(so_prime* const* primes, int nl, const int* log_bs,
                          const uint64_t* q_hats, const uint64_t* q_hat_invs,
                          uint64_t* gs) {
  int dims[64];
  so_gadget_dims(primes, nl, log_bs, dims);
  int row = 0;
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    uint64_t bi = so_import(p, (uint64_t)1 << log_bs[i]);
    uint64_t cur = so_mul(p, q_hats[i], q_hat_invs[i]);
    for (int j = 0; j < dims[i]; j++) {
      for (int l = 0; l < nl; l++) gs[(size_t)row * nl + l] = 0;
      gs[(size_t)row * nl + i] = cur;
      row++;
      cur = so_mul(p, cur, bi);
    }
  }
}

/* RnsGadget::Decompose (rns/rns_gadget.cc:108-153) with BaseDecompose
 * (gadget.h:52-79).  `a` is in coefficient form; digits receives k RNS
 * polynomials of nl limbs each (coefficient form). */
void so_gadget_decompose(so_prime* const* primes, int nl, const int* log_bs,
                         const uint64_t* a, uint64_t* digits) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  int dims[64];
  so_gadget_dims(primes, nl, log_bs, dims);
  uint64_t* cur = (uint64_t*)malloc(n * 8);
  size_t row = 0;
  for (int i = 0; i < nl; i++) {
    const so_mont* pi = &primes[i]->mp;
    const uint64_t mask = ((uint64_t)1 << log_bs[i]) - 1;
    for (size_t j = 0; j < n; j++) cur[j] = so_export(pi, a[i * n + j]);
    for (int d = 0; d < dims[i]; d++) {
      uint64_t* poly = digits + row * nl * n;
      uint64_t* own = poly + i * n;
      for (size_t j = 0; j < n; j++) {
        own[j] = so_import(pi, cur[j] & mask);
        cur[j] >>= log_bs[i];
      }
      for (int l = 0; l < nl; l++) {
        if (l == i) continue;
        so_convert_modulus_balanced(pi, &primes[l]->mp, own, poly + l * n, (int)n);
      }
      row++;
    }
  }
  free(cur);
}

/* ------------------------------------------------------------------------ */
/* Key switching (L5)                                                       */
/* ------------------------------------------------------------------------ */

/* RnsRelinKey::ApplyToRlweCiphertext (rns/rns_relinearization_key.cc:184-232).
 * comps holds degree+1 components; bit c of comp_is_ntt_mask says whether
 * component c is in NTT form (c0 and c1 must be).  Key index (i-2)*l + j. */
int so_relin_apply(so_prime* const* primes, int nl, const int* log_bs,
                   int degree, const uint64_t* comps, int comp_is_ntt_mask,
                   const uint64_t* key_as, const uint64_t* key_bs,
                   uint64_t* out_c0, uint64_t* out_c1) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  const size_t poly = (size_t)nl * n;
  int dims[64];
  const int k = so_gadget_dims(primes, nl, log_bs, dims);
  if (!(comp_is_ntt_mask & 1) || !(comp_is_ntt_mask & 2)) return -3;
  memcpy(out_c0, comps, poly * 8);
  memcpy(out_c1, comps + poly, poly * 8);
  uint64_t* ci = (uint64_t*)malloc(poly * 8);
  uint64_t* digits = (uint64_t*)malloc(poly * 8 * (size_t)k);
  for (int i = 2; i <= degree; i++) {
    memcpy(ci, comps + (size_t)i * poly, poly * 8);
    if (comp_is_ntt_mask & (1 << i)) so_poly_ntt_inv(primes, nl, ci);
    so_gadget_decompose(primes, nl, log_bs, ci, digits);
    for (int j = 0; j < k; j++) {
      uint64_t* dj = digits + (size_t)j * poly;
      so_poly_ntt_fwd(primes, nl, dj);
      const size_t index = (size_t)(i - 2) * degree + j;
      so_poly_pointwise(primes, nl, SO_FMA, out_c0, dj, key_bs + index * poly);
      so_poly_pointwise(primes, nl, SO_FMA, out_c1, dj, key_as + index * poly);
    }
  }
  free(digits);
  free(ci);
  return 0;
}

/* RnsGaloisKey::ApplyToRlweCiphertext (rns/rns_galois_key.cc:238-287). */
int so_galois_apply(so_prime* const* primes, int nl, const int* log_bs,
                    const uint64_t* c0, const uint64_t* c1, int c1_is_ntt,
                    const uint64_t* key_as, const uint64_t* key_bs,
                    uint64_t* out_c0, uint64_t* out_c1) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  const size_t poly = (size_t)nl * n;
  int dims[64];
  const int k = so_gadget_dims(primes, nl, log_bs, dims);
  uint64_t* c1c = (uint64_t*)malloc(poly * 8);
  memcpy(c1c, c1, poly * 8);
  if (c1_is_ntt) so_poly_ntt_inv(primes, nl, c1c);
  uint64_t* digits = (uint64_t*)malloc(poly * 8 * (size_t)k);
  so_gadget_decompose(primes, nl, log_bs, c1c, digits);
  memset(out_c0, 0, poly * 8);
  for (int i = 0; i < k; i++) {
    uint64_t* di = digits + (size_t)i * poly;
    so_poly_ntt_fwd(primes, nl, di);
    so_poly_pointwise(primes, nl, SO_FMA, out_c0, di, key_bs + (size_t)i * poly);
  }
  memset(out_c1, 0, poly * 8);
  for (int i = 0; i < k; i++)
    so_poly_pointwise(primes, nl, SO_FMA, out_c1, digits + (size_t)i * poly,
                      key_as + (size_t)i * poly);
  so_poly_pointwise(primes, nl, SO_ADD, out_c0, c0, NULL);
  free(digits);
  free(c1c);
  return 0;
}

/* RnsRlweCiphertext::Mul (rns/rns_ciphertext.h:501-524): out holds
 * len_a + len_b - 1 components, accumulated from zero. */
void so_tensor_product(so_prime* const* primes, int nl, int len_a,
                       const uint64_t* a, int len_b, const uint64_t* b,
                       uint64_t* out) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  const size_t poly = (size_t)nl * n;
  memset(out, 0, poly * 8 * (size_t)(len_a + len_b - 1));
  for (int i = 0; i < len_a; i++)
    for (int j = 0; j < len_b; j++)
      so_poly_pointwise(primes, nl, SO_FMA, out + (size_t)(i + j) * poly,
                        a + (size_t)i * poly, b + (size_t)j * poly);
}

/* ------------------------------------------------------------------------ */
/* LazyRnsPolynomial (rns/lazy_rns_polynomial.{h,cc})                       */
/* ------------------------------------------------------------------------ */

/* ComputeLogMaximalLevel for one modulus (rns/lazy_rns_polynomial.h:156-181). */
int so_lazy_log_max_level(const so_mont* p) {
  int lvl = 64 - 2;
  int lq = 128 - 1 - 2 * (int)p->log_modulus;
  return lq < lvl ? lq : lvl;
}

/* CreateZero, num_terms FusedMulAddInPlace (with Refresh at the maximal
 * level), then Export (rns/lazy_rns_polynomial.h:84-129,
 * rns/lazy_rns_polynomial.cc:134-151, rns/rns_polynomial_hwy.cc:122-156).
 * as, bs: num_terms RNS polynomials each. */
void so_lazy_fma(so_prime* const* primes, int nl, int num_terms,
                 const uint64_t* as, const uint64_t* bs, uint64_t* out) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  const size_t poly = (size_t)nl * n;
  int log_max = 62;
  for (int i = 0; i < nl; i++) {
    int l = so_lazy_log_max_level(&primes[i]->mp);
    if (l < log_max) log_max = l;
  }
  const uint64_t max_level = (uint64_t)1 << log_max;
  u128* acc = (u128*)calloc(poly, sizeof(u128));
  uint64_t level = 0;
  for (int t = 0; t < num_terms; t++) {
    if (level == max_level) {
      for (int i = 0; i < nl; i++)
        for (size_t j = 0; j < n; j++) {
          u128 v = acc[i * n + j];
          acc[i * n + j] = so_barrett_reduce_big(&primes[i]->mp, (uint64_t)v,
                                                 (uint64_t)(v >> 64));
        }
      level = 1;
    }
    const uint64_t* a = as + (size_t)t * poly;
    const uint64_t* b = bs + (size_t)t * poly;
    for (size_t j = 0; j < poly; j++) acc[j] += (u128)a[j] * b[j];
    level++;
  }
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    for (size_t j = 0; j < n; j++) {
      u128 v = acc[i * n + j];
      out[i * n + j] = so_export(p, so_barrett_reduce_big(p, (uint64_t)v,
                                                          (uint64_t)(v >> 64)));
    }
  }
  free(acc);
}

/* LazyRnsPolynomial over mixed updates (rns/lazy_rns_polynomial.h:44-115,
 * rns/lazy_rns_polynomial.cc:89-254): starting from CreateZero, term t adds
 * the integer product a b (kind 0, FusedMulAddInPlace), (a + b) c (kind 1,
 * FusedMulSumAddInPlace) or (a + q_i - b) c (kind 2,
 * FusedMulDifferenceAddInPlace) of the Montgomery representations to a
 * 128-bit sum, refreshing (BarrettReduceBigInt) when the level reaches
 * 2^log_max_level; Export reduces and applies ExportInt. */
void so_lazy_mixed(so_prime* const* primes, int nl, int num_terms, const int* kinds,
                   const uint64_t* as, const uint64_t* bs, const uint64_t* cs, uint64_t* out) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  const size_t poly = (size_t)nl * n;
  int log_max = 62;
  for (int i = 0; i < nl; i++) {
    int l = so_lazy_log_max_level(&primes[i]->mp);
    if (l < log_max) log_max = l;
  }
  const uint64_t max_level = (uint64_t)1 << log_max;
  u128* acc = (u128*)calloc(poly, sizeof(u128));
  uint64_t level = 0;
  for (int t = 0; t < num_terms; t++) {
    if (level == max_level) {
      for (int i = 0; i < nl; i++)
        for (size_t j = 0; j < n; j++) {
          u128 v = acc[i * n + j];
          acc[i * n + j] = so_barrett_reduce_big(&primes[i]->mp, (uint64_t)v,
                                                 (uint64_t)(v >> 64));
        }
      level = 1;
    }
    const uint64_t* a = as + (size_t)t * poly;
    const uint64_t* b = bs + (size_t)t * poly;
    const uint64_t* c = cs + (size_t)t * poly;
    for (int i = 0; i < nl; i++) {
      const uint64_t q = primes[i]->mp.modulus;
      for (size_t j = i * n; j < (i + 1) * n; j++) {
        if (kinds[t] == 0)
          acc[j] += (u128)a[j] * b[j];
        else if (kinds[t] == 1)
          acc[j] += (u128)(a[j] + b[j]) * c[j];
        else
          acc[j] += (u128)(a[j] + q - b[j]) * c[j];
      }
    }
    level++;
  }
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    for (size_t j = 0; j < n; j++) {
      u128 v = acc[i * n + j];
      out[i * n + j] = so_export(p, so_barrett_reduce_big(p, (uint64_t)v,
                                                          (uint64_t)(v >> 64)));
    }
  }
  free(acc);
}

/* ------------------------------------------------------------------------ */
/* Deterministic inputs                                                      */
/* ------------------------------------------------------------------------ */

uint64_t so_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Uniform residues in [0, q_i) stored directly as Montgomery representations,
 * by rejection on log_modulus bits as ImportRandom does
 * (montgomery.h:260-273), with SplitMix64 as the bit source. */
void so_fill_uniform(so_prime* const* primes, int nl, uint64_t seed,
                     uint64_t* x) {
  uint64_t st = seed;
  for (int i = 0; i < nl; i++) {
    const so_mont* p = &primes[i]->mp;
    const size_t n = (size_t)1 << primes[i]->log_n;
    const uint64_t mask = (p->log_modulus >= 64) ? ~0ull
                                                 : (((uint64_t)1 << p->log_modulus) - 1);
    for (size_t j = 0; j < n; j++) {
      uint64_t v;
      do { v = so_splitmix64(&st) & mask; } while (v >= p->modulus);
      x[i * n + j] = v;
    }
  }
}

double so_now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* L8: wire formats and the ChaCha PRNG                                      */
/* ------------------------------------------------------------------------ */

static uint64_t so_low_mask(int bits) { return bits >= 64 ? ~0ull : (((uint64_t)1 << bits) - 1); }

/* TranscribeBits<InputInt, OutputInt> (transcription.h:60-150): `in` holds
 * input integers of in_type_bits (8 or 64) bits, of which the low
 * input_bits_per_int are used; output integers receive
 * output_bits_per_int bits each.  Returns the number of output integers, or
 * a negative error (the reference's InvalidArgument cases). */
long long so_transcribe_bits(const uint64_t* in, size_t in_len, int in_type_bits,
                             long long input_bit_length, int input_bits_per_int,
                             int output_bits_per_int, int out_type_bits, uint64_t* out,
                             size_t out_cap) {
  if (input_bits_per_int > in_type_bits) return -1;
  if (output_bits_per_int > out_type_bits) return -2;
  if (input_bit_length < 0) return -3;
  if (input_bit_length == 0) return in_len == 0 ? 0 : -4;
  const size_t input_chunks =
      (size_t)((input_bit_length + input_bits_per_int - 1) / input_bits_per_int);
  if (in_len < input_chunks) return -5;
  const size_t output_chunks =
      (size_t)((input_bit_length + output_bits_per_int - 1) / output_bits_per_int);
  if (out_cap < output_chunks) return -6;
  memset(out, 0, output_chunks * sizeof(uint64_t));
  const uint64_t in_mask = so_low_mask(in_type_bits);
  long long remaining = input_bit_length;
  size_t op = 0;
  int size_out = (int)(remaining < output_bits_per_int ? remaining : output_bits_per_int);
  int needed = size_out;
  for (size_t i = 0; i < input_chunks; i++) {
    int nin = (int)(input_bits_per_int < remaining ? input_bits_per_int : remaining);
    /* The chunk's bits moved to the top of an input-width integer. */
    const uint64_t input = (in[i] << (in_type_bits - nin)) & in_mask;
    while (nin > 0) {
      if (needed == 0) {
        op++;
        size_out = (int)(remaining < output_bits_per_int ? remaining : output_bits_per_int);
        needed = size_out;
      }
      const int nproc = nin < needed ? nin : needed;
      const uint64_t bits_left = (input << (nin - nproc)) & in_mask;
      const uint64_t mask = bits_left >> (in_type_bits - nproc);
      out[op] |= (mask << (size_out - needed)) & so_low_mask(out_type_bits);
      nin -= nproc;
      needed -= nproc;
    }
    remaining -= input_bits_per_int;
  }
  return (long long)output_chunks;
}

/* MontgomeryInt<Uint64>::SerializeVector (montgomery.cc:143-167): the
 * representations' low log_modulus bits, packed into bytes.  Returns the
 * byte count, or a negative error. */
long long so_serialize_vector(const so_mont* p, const uint64_t* coeffs, size_t count,
                              uint8_t* out, size_t out_cap) {
  if (count > ((size_t)1 << 15)) return -10;  /* kMaxNumCoeffs (constants.h:75) */
  if (count == 0) return -11;                 /* "Cannot serialize an empty vector." */
  const int b = (int)p->log_modulus;
  const size_t nbytes = (count * (size_t)b + 7) / 8;
  if (out_cap < nbytes) return -6;
  uint64_t* tmp = (uint64_t*)malloc(nbytes * sizeof(uint64_t));
  const long long r = so_transcribe_bits(coeffs, count, 64, (long long)count * b, b, 8, 8, tmp,
                                         nbytes);
  for (long long i = 0; i < r; i++) out[i] = (uint8_t)tmp[i];
  free(tmp);
  return r;
}

/* MontgomeryInt<Uint64>::DeserializeVector (montgomery.cc:184-219): the
 * coefficients are taken as representations without reduction. */
int so_deserialize_vector(const so_mont* p, int num_coeffs, const uint8_t* in, size_t in_len,
                          uint64_t* coeffs) {
  if (num_coeffs < 0) return -12;
  if (num_coeffs > (1 << 15)) return -10;
  const int b = (int)p->log_modulus;
  uint64_t* tmp = (uint64_t*)malloc((in_len ? in_len : 1) * sizeof(uint64_t));
  for (size_t i = 0; i < in_len; i++) tmp[i] = in[i];
  const long long r = so_transcribe_bits(tmp, in_len, 8, (long long)b * num_coeffs, 8, b, 64,
                                         coeffs, (size_t)num_coeffs);
  free(tmp);
  if (r < 0) return (int)r;
  if (r < num_coeffs) return -13; /* "Given serialization is invalid." */
  return 0;
}

/* ChaCha20 block function of RFC 7539 Sec. 2.3 (used by the reference through
 * BoringSSL's CRYPTO_chacha_20, prng/chacha_prng_util.cc:36-60). */
#define SO_ROTL32(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define SO_QR(a, b, c, d)            \
  a += b; d ^= a; d = SO_ROTL32(d, 16); \
  c += d; b ^= c; b = SO_ROTL32(b, 12); \
  a += b; d ^= a; d = SO_ROTL32(d, 8);  \
  c += d; b ^= c; b = SO_ROTL32(b, 7);
static uint32_t so_le32(const uint8_t* p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
void so_chacha20_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                       uint8_t out[64]) {
  uint32_t s[16], x[16];
  s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
  for (int i = 0; i < 8; i++) s[4 + i] = so_le32(key + 4 * i);
  s[12] = counter;
  for (int i = 0; i < 3; i++) s[13 + i] = so_le32(nonce + 4 * i);
  memcpy(x, s, sizeof(s));
  for (int r = 0; r < 10; r++) {
    SO_QR(x[0], x[4], x[8], x[12]) SO_QR(x[1], x[5], x[9], x[13])
    SO_QR(x[2], x[6], x[10], x[14]) SO_QR(x[3], x[7], x[11], x[15])
    SO_QR(x[0], x[5], x[10], x[15]) SO_QR(x[1], x[6], x[11], x[12])
    SO_QR(x[2], x[7], x[8], x[13]) SO_QR(x[3], x[4], x[9], x[14])
  }
  for (int i = 0; i < 16; i++) {
    const uint32_t v = x[i] + s[i];
    out[4 * i] = (uint8_t)v; out[4 * i + 1] = (uint8_t)(v >> 8);
    out[4 * i + 2] = (uint8_t)(v >> 16); out[4 * i + 3] = (uint8_t)(v >> 24);
  }
}

/* ChaChaPrngResalt (prng/chacha_prng_util.cc:27-60): the buffer is the
 * ChaCha20 keystream (block counter from 0) under nonce "salt" followed by
 * the 64-bit little-endian salt counter, kChaChaOutputBytes = 255 * 32 long. */
static void so_chacha_resalt(so_chacha_prng* g) {
  uint8_t nonce[12] = {'s', 'a', 'l', 't'};
  uint64_t c = (uint64_t)g->salt_counter;
  for (int i = 0; i < 8; i++) { nonce[4 + i] = (uint8_t)(c & 0xff); c >>= 8; }
  for (int blk = 0; blk * 64 < SO_CHACHA_OUTPUT_BYTES; blk++) {
    uint8_t ks[64];
    so_chacha20_block(g->key, (uint32_t)blk, nonce, ks);
    const int take = SO_CHACHA_OUTPUT_BYTES - blk * 64 < 64 ? SO_CHACHA_OUTPUT_BYTES - blk * 64 : 64;
    memcpy(g->buffer + blk * 64, ks, (size_t)take);
  }
  g->salt_counter++;
  g->position = 0;
}

/* SingleThreadChaChaPrng::Create (prng/single_thread_chacha_prng.cc:36-52):
 * a 32-byte key, the first buffer drawn at once. */
int so_chacha_prng_init(so_chacha_prng* g, const uint8_t* key, size_t key_len) {
  if (key_len != 32) return -1;
  memcpy(g->key, key, 32);
  g->salt_counter = 0;
  so_chacha_resalt(g);
  return 0;
}

/* ChaChaPrngRand8 / ChaChaPrngRand64 (prng/chacha_prng_util.cc:76-104):
 * bytes in order, eight of them little-endian for 64 bits. */
uint8_t so_chacha_rand8(so_chacha_prng* g) {
  if (g->position >= SO_CHACHA_OUTPUT_BYTES) so_chacha_resalt(g);
  return g->buffer[g->position++];
}
uint64_t so_chacha_rand64(so_chacha_prng* g) {
  uint64_t r = 0;
  for (int i = 0; i < 8; i++) r += (uint64_t)so_chacha_rand8(g) << (8 * i);
  return r;
}

/* MontgomeryInt::GenerateRandomInt (montgomery.h:631-660) and ImportRandom
 * (montgomery.h:260-273): log_modulus random bits (64-bit draws, a final
 * 8-bit draw for the last <= 8 bits), rejected until below the modulus and
 * stored directly as the representation. */
uint64_t so_import_random_chacha(const so_mont* p, so_chacha_prng* g) {
  for (;;) {
    int bits = (int)p->log_modulus;
    uint64_t r = 0;
    while (bits > 0) {
      if (bits <= 8) {
        const uint64_t v = so_chacha_rand8(g);
        r = (r << bits) + (v & so_low_mask(bits));
        break;
      }
      const uint64_t v = so_chacha_rand64(g);
      const int take = bits < 64 ? bits : 64;
      r = (take >= 64 ? 0 : (r << take)) + (v & so_low_mask(take));
      bits -= take;
    }
    if (r < p->modulus) return r;
  }
}

/* RnsPolynomial::SampleUniform (rns/rns_polynomial.h:640-667): limb by limb,
 * coefficient by coefficient, from one PRNG. */
void so_sample_uniform_chacha(so_prime* const* primes, int nl, so_chacha_prng* g, uint64_t* x) {
  for (int i = 0; i < nl; i++) {
    const size_t n = (size_t)1 << primes[i]->log_n;
    for (size_t j = 0; j < n; j++) x[i * n + j] = so_import_random_chacha(&primes[i]->mp, g);
  }
}

/* ------------------------------------------------------------------------ */
/* L9: BFV-side RNS operations (rns/rns_polynomial.cc:349-852)               */
/* ------------------------------------------------------------------------ */

/* ConvertModulus (modulus_conversion.h:140-157): export mod q, import mod p. */
static void so_convert_modulus(const so_mont* pq, const so_mont* pp, const uint64_t* in,
                               uint64_t* out, int n) {
  if (pq->modulus == pp->modulus) {
    memmove(out, in, sizeof(uint64_t) * (size_t)n);
    return;
  }
  for (int i = 0; i < n; i++) out[i] = so_import(pp, so_export(pq, in[i]));
}

/* RnsPolynomial::ModReduceMsb (rns/rns_polynomial.cc:349-417): on return
 * the first (nl - 1) limbs hold (a - [a]_{q_l}) / q_l. */
int so_poly_mod_reduce_msb(so_prime* const* primes, int nl, const uint64_t* ql_inv, int is_ntt,
                           uint64_t* x) {
  if (nl <= 1) return -1;
  const size_t n = (size_t)1 << primes[0]->log_n;
  const int nr = nl - 1;
  const so_prime* pl = primes[nl - 1];
  uint64_t* delta_ql = (uint64_t*)malloc(n * 8);
  memcpy(delta_ql, x + (size_t)nr * n, n * 8);
  uint64_t* delta = (uint64_t*)malloc(n * 8 * (size_t)nr);
  if (is_ntt) {
    so_ntt_inv(pl, delta_ql);
    for (int i = 0; i < nr; i++)
      so_convert_modulus_balanced(&pl->mp, &primes[i]->mp, delta_ql, delta + i * n, (int)n);
    so_poly_ntt_fwd(primes, nr, delta);
    so_poly_pointwise(primes, nr, SO_SUB, x, delta, NULL);
    so_poly_mul_scalar(primes, nr, x, ql_inv);
  } else {
    for (int i = 0; i < nr; i++) {
      const so_mont* mi = &primes[i]->mp;
      uint64_t* di = delta + i * n;
      so_convert_modulus_balanced(&pl->mp, mi, delta_ql, di, (int)n);
      for (size_t j = 0; j < n; j++) x[i * n + j] = so_sub(mi, x[i * n + j], di[j]);
      for (size_t j = 0; j < n; j++) x[i * n + j] = so_mul(mi, x[i * n + j], ql_inv[i]);
    }
  }
  free(delta);
  free(delta_ql);
  return 0;
}

/* y_i = a_i * q_hat_inv_i mod q_i for every limb, and u = round(sum_k
 * export(y_k) / q_k) per coefficient, in the reference's order of floating
 * point operations (rns/rns_polynomial.cc:451-483). */
static void so_scaled_and_u(so_prime* const* in_p, int L, const uint64_t* q_hat_inv,
                            const uint64_t* pre_scale, const uint64_t* in, uint64_t* y,
                            double* us) {
  const size_t n = (size_t)1 << in_p[0]->log_n;
  for (size_t k = 0; k < n; k++) us[k] = 0.0;
  for (int i = 0; i < L; i++) {
    const so_mont* mi = &in_p[i]->mp;
    for (size_t k = 0; k < n; k++) {
      uint64_t v = in[i * n + k];
      if (pre_scale) v = so_mul(mi, v, pre_scale[i]);
      y[i * n + k] = so_mul(mi, v, q_hat_inv[i]);
    }
    const double qd = (double)mi->modulus;
    for (size_t k = 0; k < n; k++) {
      double s = (double)so_export(mi, y[i * n + k]);
      s /= qd;
      us[k] += s;
    }
  }
}

/* RnsPolynomial::SwitchRnsBasis (rns/rns_polynomial.cc:419-517): exact
 * conversion of a coefficient-form polynomial mod Q to mod P.
 * q_hat_mod_ps[i * K + j] = (Q/q_i) mod p_j, q_mod_ps[j] = Q mod p_j. */
void so_switch_rns_basis(so_prime* const* in_p, int L, so_prime* const* out_p, int K,
                         const uint64_t* q_hat_inv, const uint64_t* q_hat_mod_ps,
                         const uint64_t* q_mod_ps, const uint64_t* in, uint64_t* out) {
  const size_t n = (size_t)1 << in_p[0]->log_n;
  uint64_t* y = (uint64_t*)malloc(n * 8 * (size_t)L);
  double* us = (double*)malloc(n * sizeof(double));
  uint64_t* tmp = (uint64_t*)malloc(n * 8);
  so_scaled_and_u(in_p, L, q_hat_inv, NULL, in, y, us);
  memset(out, 0, n * 8 * (size_t)K);
  for (int i = 0; i < L; i++) {
    for (int j = 0; j < K; j++) {
      const so_mont* mj = &out_p[j]->mp;
      so_convert_modulus(&in_p[i]->mp, mj, y + i * n, tmp, (int)n);
      for (size_t k = 0; k < n; k++) {
        tmp[k] = so_mul(mj, tmp[k], q_hat_mod_ps[i * K + j]);
        out[j * n + k] = so_add(mj, out[j * n + k], tmp[k]);
      }
    }
  }
  for (int j = 0; j < K; j++) {
    const so_mont* mj = &out_p[j]->mp;
    for (size_t k = 0; k < n; k++) {
      const uint64_t u = (uint64_t)round(us[k]);
      const uint64_t c = so_mul(mj, so_import(mj, u), q_mod_ps[j]);
      out[j * n + k] = so_sub(mj, out[j * n + k], c);
    }
  }
  free(tmp);
  free(us);
  free(y);
}

/* RnsPolynomial::ScaleAndSwitchRnsBasis (rns/rns_polynomial.cc:519-636):
 * round(P * a / Q) mod P for a coefficient-form polynomial mod Q.
 * q_inv_mod_ps[j * L + i] = q_i^-1 mod p_j, p_mod_qs[i] = P mod q_i. */
void so_scale_and_switch_rns_basis(so_prime* const* in_p, int L, so_prime* const* out_p, int K,
                                   const uint64_t* q_hat_inv, const uint64_t* q_inv_mod_ps,
                                   const uint64_t* p_mod_qs, const uint64_t* in, uint64_t* out) {
  const size_t n = (size_t)1 << in_p[0]->log_n;
  uint64_t* y = (uint64_t*)malloc(n * 8 * (size_t)L);
  double* us = (double*)malloc(n * sizeof(double));
  uint64_t* tmp = (uint64_t*)malloc(n * 8);
  so_scaled_and_u(in_p, L, q_hat_inv, p_mod_qs, in, y, us);
  memset(out, 0, n * 8 * (size_t)K);
  for (int i = 0; i < L; i++) {
    for (int j = 0; j < K; j++) {
      const so_mont* mj = &out_p[j]->mp;
      so_convert_modulus(&in_p[i]->mp, mj, y + i * n, tmp, (int)n);
      const uint64_t neg = so_neg(mj, q_inv_mod_ps[j * L + i]);
      for (size_t k = 0; k < n; k++) {
        tmp[k] = so_mul(mj, tmp[k], neg);
        out[j * n + k] = so_add(mj, out[j * n + k], tmp[k]);
      }
    }
  }
  for (int j = 0; j < K; j++) {
    const so_mont* mj = &out_p[j]->mp;
    for (size_t k = 0; k < n; k++) {
      const uint64_t u = (uint64_t)round(us[k]);
      out[j * n + k] = so_add(mj, out[j * n + k], so_import(mj, u));
    }
  }
  free(tmp);
  free(us);
  free(y);
}

/* RnsPolynomial::ApproxSwitchRnsBasis (rns/rns_polynomial.cc:638-701): the
 * fast basis conversion sum_i [a_i (Q/q_i)^-1]_{q_i} (Q/q_i) mod p_j, the
 * bracket lifted balanced or not; input in either form, output in
 * coefficient form. */
void so_approx_switch_rns_basis(so_prime* const* in_p, int L, so_prime* const* out_p, int K,
                                const uint64_t* q_hat_inv, const uint64_t* q_hat_mod_ps,
                                int is_ntt, int balanced, const uint64_t* in, uint64_t* out) {
  const size_t n = (size_t)1 << in_p[0]->log_n;
  uint64_t* a = (uint64_t*)malloc(n * 8);
  uint64_t* tmp = (uint64_t*)malloc(n * 8);
  memset(out, 0, n * 8 * (size_t)K);
  for (int i = 0; i < L; i++) {
    const so_mont* mi = &in_p[i]->mp;
    for (size_t k = 0; k < n; k++) a[k] = so_mul(mi, in[i * n + k], q_hat_inv[i]);
    if (is_ntt) so_ntt_inv(in_p[i], a);
    for (int j = 0; j < K; j++) {
      const so_mont* mj = &out_p[j]->mp;
      if (balanced) so_convert_modulus_balanced(mi, mj, a, tmp, (int)n);
      else so_convert_modulus(mi, mj, a, tmp, (int)n);
      for (size_t k = 0; k < n; k++) {
        tmp[k] = so_mul(mj, tmp[k], q_hat_mod_ps[i * K + j]);
        out[j * n + k] = so_add(mj, out[j * n + k], tmp[k]);
      }
    }
  }
  free(tmp);
  free(a);
}
