/* Vectorised CPU restatements for the host baseline (test infrastructure,
 * like the rest of this directory: timed by the benchmark's CPU leg and
 * checked word for word against the scalar restatements in tests; never
 * part of the GPU library).
 *
 * so_hwy_ntt_fwd restates the reference's Highway forward transform,
 * IterativeCooleyTukeyHwy (dft_transformations_hwy.cc:64-158), for vectors
 * of eight 64-bit lanes (the AVX-512 width): every layer whose half-length
 * is at least eight runs Harvey's lazy butterflies on whole vectors (Shoup
 * product t = c v - hi(c' v) q in [0, 2q), u reduced to [0, 2q), outputs in
 * [0, 4q), the last vector layer reduced to [0, 2q)), and the last three
 * layers run the scalar butterflies with alternating lazy additions.  The
 * output is the canonical transform, equal to so_ntt_fwd's.
 *
 * so_hwy_relin_apply is RnsRelinKey::ApplyToRlweCiphertext
 * (rns/rns_relinearization_key.cc:184-232) as so_relin_apply computes it,
 * with that transform for the digits and FusedMulAddInPlace
 * (montgomery.h:446-465, REDC(R acc + a b)) on whole vectors.
 *
 * 64 x 64 -> 128-bit products are formed from 32-bit halves (four widening
 * 32-bit multiplies per high word), as Highway's MulHigh does on x86.  The
 * functions are compiled for AVX-512, AVX2 and baseline x86-64 and the
 * loader picks the widest the host supports. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "shell_oracle.h"

typedef uint64_t vu64 __attribute__((vector_size(64)));
#define VL 8
#define LOG_VL 3

#if defined(__x86_64__) && !defined(SO_NO_CLONES)
#define SO_CLONES __attribute__((target_clones("avx512f", "avx2", "default")))
#else
#define SO_CLONES
#endif

static inline vu64 vset(uint64_t x) { return (vu64){x, x, x, x, x, x, x, x}; }
static inline vu64 vload(const uint64_t* p) {
  vu64 v;
  memcpy(&v, p, sizeof v);
  return v;
}
static inline void vstore(uint64_t* p, vu64 v) { memcpy(p, &v, sizeof v); }

/* High 64 bits of a * b per lane. */
static inline vu64 vmulhi(vu64 a, vu64 b) {
  const vu64 m = vset(0xffffffffull);
  const vu64 a0 = a & m, a1 = a >> 32, b0 = b & m, b1 = b >> 32;
  const vu64 p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
  const vu64 mid = (p00 >> 32) + (p01 & m) + (p10 & m);
  return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

/* x >= m ? x - m : x per lane. */
static inline vu64 vcsub(vu64 x, vu64 m) { return x - (m & (vu64)(x >= m)); }

SO_CLONES
void so_hwy_ntt_fwd(const so_prime* pr, uint64_t* x) {
  const so_mont* p = &pr->mp;
  const uint64_t q = p->modulus;
  const int n = 1 << pr->log_n;
  const vu64 neg_q = vset(0 - q), two_q = vset(2 * q);
  int index_psi = 1;
  int i = pr->log_n - 1;
  for (; i >= LOG_VL; i--) {
    const int half_m = 1 << i;
    const int m = half_m << 1;
    for (int k = 0; k < n; k += m) {
      const vu64 c = vset(pr->psis_bitrev_c[index_psi]);
      const vu64 cb = vset(pr->psis_bitrev_cb[index_psi]);
      for (int j = 0; j < half_m; j += VL) {
        uint64_t* ptr = x + k + j;
        const vu64 v = vload(ptr + half_m);
        const vu64 t = c * v + vmulhi(cb, v) * neg_q; /* [0, 2q) */
        vu64 u = vcsub(vload(ptr), two_q);            /* [0, 2q) */
        vu64 w = u + (two_q - t);                     /* [0, 4q) */
        u = u + t;
        if (i == LOG_VL) {
          u = vcsub(u, two_q);
          w = vcsub(w, two_q);
        }
        vstore(ptr, u);
        vstore(ptr + half_m, w);
      }
      index_psi++;
    }
  }
  int lazy = 0;
  for (; i >= 0; i--) {
    const int half_m = 1 << i;
    const int m = half_m << 1;
    for (int k = 0; k < n; k += m) {
      const uint64_t c = pr->psis_bitrev_c[index_psi];
      const uint64_t cb = pr->psis_bitrev_cb[index_psi];
      for (int j = 0; j < half_m; j++) {
        const uint64_t t = so_mul_const(p, x[k + j + half_m], c, cb);
        const uint64_t u = x[k + j];
        if (i != 0 && lazy) {
          x[k + j] = u + t;
          x[k + j + half_m] = u + (q - t);
        } else {
          x[k + j] = so_add(p, u, t);
          x[k + j + half_m] = so_sub(p, u, t);
        }
      }
      index_psi++;
    }
    lazy = !lazy;
  }
}

void so_hwy_poly_ntt_fwd(so_prime* const* primes, int nl, uint64_t* x) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  for (int i = 0; i < nl; i++) so_hwy_ntt_fwd(primes[i], x + (size_t)i * n);
}

/* acc <- REDC(R acc + a b) = acc + a b R^-1 mod q, canonical, n words
 * (n a multiple of 8). */
SO_CLONES
static void hwy_fma_limb(const so_mont* p, uint64_t* acc, const uint64_t* a, const uint64_t* b,
                         size_t n) {
  const vu64 q = vset(p->modulus), r = vset(p->r_mod_modulus), inv = vset(p->inv_modulus);
  for (size_t j = 0; j < n; j += VL) {
    const vu64 x = vload(acc + j), va = vload(a + j), vb = vload(b + j);
    /* t = r x + a b as (hi, lo) */
    const vu64 l1 = r * x, h1 = vmulhi(r, x);
    const vu64 l2 = va * vb, h2 = vmulhi(va, vb);
    vu64 lo = l1 + l2;
    vu64 hi = h1 + h2 + (vset(1) & (vu64)(lo < l1));
    /* + q u with u = lo * inv: the low word becomes zero */
    const vu64 u = lo * inv;
    const vu64 l3 = q * u, h3 = vmulhi(q, u);
    const vu64 lo2 = lo + l3;
    hi = hi + h3 + (vset(1) & (vu64)(lo2 < lo));
    vstore(acc + j, vcsub(hi, q));
  }
}

int so_hwy_relin_apply(so_prime* const* primes, int nl, const int* log_bs, int degree,
                       const uint64_t* comps, int comp_is_ntt_mask, const uint64_t* key_as,
                       const uint64_t* key_bs, uint64_t* out_c0, uint64_t* out_c1) {
  const size_t n = (size_t)1 << primes[0]->log_n;
  const size_t poly = (size_t)nl * n;
  if (n % VL) return -4;
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
      so_hwy_poly_ntt_fwd(primes, nl, dj);
      const size_t index = (size_t)(i - 2) * degree + j;
      for (int l = 0; l < nl; l++) {
        const size_t off = (size_t)l * n;
        hwy_fma_limb(&primes[l]->mp, out_c0 + off, dj + off, key_bs + index * poly + off, n);
        hwy_fma_limb(&primes[l]->mp, out_c1 + off, dj + off, key_as + index * poly + off, n);
      }
    }
  }
  free(digits);
  free(ci);
  return 0;
}
