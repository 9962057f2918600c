/*
 * shell_oracle.h -- CPU restatement of SHELL's RNS polynomial hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP path is tested against this oracle:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libshellgpu.so) never links or calls it.
 *
 * Every function restates one reference routine, cited as file:line of
 * google/shell-encryption (paths relative to the repository root).  Values are
 * MontgomeryInt<Uint64> representations (x * 2^64 mod q), exactly as the
 * reference stores them.  Polynomials are limb-major arrays: limb i of an RNS
 * polynomial with n coefficients occupies words [i*n, (i+1)*n).
 */
#ifndef SHELL_ORACLE_H_
#define SHELL_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MontgomeryIntParams<Uint64> (shell_encryption/montgomery.h:77-192). */
typedef struct so_mont {
  uint64_t modulus;
  unsigned log_modulus;
  uint64_t r_mod_modulus;
  uint64_t r_mod_modulus_barrett;
  uint64_t inv_modulus;
  uint64_t inv_r;
  uint64_t inv_r_barrett;
  uint64_t barrett_numerator;        /* floor(2^64 / q) */
  uint64_t barrett_numerator_big_lo; /* floor(2^127 / q), low word */
  uint64_t barrett_numerator_big_hi; /* floor(2^127 / q), high word */
} so_mont;

/* PrimeModulus<MontgomeryInt<Uint64>> = params + NttParameters
 * (shell_encryption/rns/rns_modulus.h:30-42, ntt_parameters.h:160-179). */
typedef struct so_prime {
  so_mont mp;
  int log_n;
  uint64_t psi;                 /* primitive 2n-th root (Montgomery rep) */
  uint64_t n_inv;               /* n^-1 (Montgomery rep) */
  uint64_t* psis_bitrev;        /* n entries, Montgomery reps */
  uint64_t* psis_bitrev_c;      /* GetConstant().first  */
  uint64_t* psis_bitrev_cb;     /* GetConstant().second */
  uint64_t* psis_inv_bitrev;
  uint64_t* psis_inv_bitrev_c;
  uint64_t* psis_inv_bitrev_cb;
  uint32_t* bitrevs;
} so_prime;

/* ---- L1: Montgomery arithmetic (montgomery.h / montgomery.cc) ---- */
int so_mont_init(uint64_t q, so_mont* p);
uint64_t so_import(const so_mont* p, uint64_t x);
uint64_t so_export(const so_mont* p, uint64_t x);
uint64_t so_barrett_reduce(const so_mont* p, uint64_t x);
uint64_t so_barrett_reduce_big(const so_mont* p, uint64_t lo, uint64_t hi);
uint64_t so_mul(const so_mont* p, uint64_t a, uint64_t b);
uint64_t so_mul_const(const so_mont* p, uint64_t x, uint64_t c, uint64_t cb);
uint64_t so_add(const so_mont* p, uint64_t a, uint64_t b);
uint64_t so_sub(const so_mont* p, uint64_t a, uint64_t b);
uint64_t so_neg(const so_mont* p, uint64_t a);
uint64_t so_fma(const so_mont* p, uint64_t acc, uint64_t a, uint64_t b);
uint64_t so_fma_const(const so_mont* p, uint64_t acc, uint64_t a, uint64_t c,
                      uint64_t cb);
void so_get_constant(const so_mont* p, uint64_t w, uint64_t* c, uint64_t* cb);
uint64_t so_modexp(const so_mont* p, uint64_t base, uint64_t e);
uint64_t so_inverse(const so_mont* p, uint64_t a);

/* ---- L2: NTT parameters and transforms ---- */
so_prime* so_prime_create(uint64_t q, int log_n, int* status);
void so_prime_destroy(so_prime* pr);
const so_mont* so_prime_mont(const so_prime* pr);
void so_prime_tables(const so_prime* pr, uint64_t* out6n, uint32_t* bitrevs,
                     uint64_t* psi_ninv);
void so_ntt_fwd(const so_prime* pr, uint64_t* x);
void so_ntt_inv(const so_prime* pr, uint64_t* x);

/* ---- L3: RNS polynomial operations over `nl` limbs ---- */
typedef enum {
  SO_ADD = 0, SO_SUB = 1, SO_MUL = 2, SO_NEG = 3, SO_FMA = 4
} so_pointwise_op;
void so_poly_pointwise(so_prime* const* primes, int nl, int op, uint64_t* x,
                       const uint64_t* a, const uint64_t* b);
void so_poly_mul_scalar(so_prime* const* primes, int nl, uint64_t* x,
                        const uint64_t* scalars);
void so_poly_mul_int(so_prime* const* primes, int nl, uint64_t* x,
                     uint64_t scalar);
void so_poly_ntt_fwd(so_prime* const* primes, int nl, uint64_t* x);
void so_poly_ntt_inv(so_prime* const* primes, int nl, uint64_t* x);
int so_poly_substitute(so_prime* const* primes, int nl, int power,
                       const uint64_t* in, uint64_t* out);
void so_convert_modulus_balanced(const so_mont* pq, const so_mont* pp,
                                 const uint64_t* in, uint64_t* out, int n);
int so_poly_mod_reduce_lsb(so_prime* const* primes, int nl, uint64_t t,
                           const uint64_t* ql_inv, int is_ntt, uint64_t* x);

/* RnsGadget: dims[i] = GadgetSize(log_bs[i], q_i); returns k = sum dims. */
int so_gadget_dims(so_prime* const* primes, int nl, const int* log_bs,
                   int* dims);
void so_gadget_components(so_prime* const* primes, int nl, const int* log_bs,
                          const uint64_t* q_hats, const uint64_t* q_hat_invs,
                          uint64_t* gs);
void so_gadget_decompose(so_prime* const* primes, int nl, const int* log_bs,
                         const uint64_t* a, uint64_t* digits);

/* Key switching. Components are limb-major RNS polynomials of nl*n words. */
int so_relin_apply(so_prime* const* primes, int nl, const int* log_bs,
                   int degree, const uint64_t* comps, int comp_is_ntt_mask,
                   const uint64_t* key_as, const uint64_t* key_bs,
                   uint64_t* out_c0, uint64_t* out_c1);
int so_galois_apply(so_prime* const* primes, int nl, const int* log_bs,
                    const uint64_t* c0, const uint64_t* c1, int c1_is_ntt,
                    const uint64_t* key_as, const uint64_t* key_bs,
                    uint64_t* out_c0, uint64_t* out_c1);
void so_tensor_product(so_prime* const* primes, int nl, int len_a,
                       const uint64_t* a, int len_b, const uint64_t* b,
                       uint64_t* out);

/* LazyRnsPolynomial fused multiply-add and export. acc holds 2 words
 * (lo, hi) per coefficient. */
int so_lazy_log_max_level(const so_mont* p);
void so_lazy_fma(so_prime* const* primes, int nl, int num_terms,
                 const uint64_t* as, const uint64_t* bs, uint64_t* out);
void so_lazy_mixed(so_prime* const* primes, int nl, int num_terms, const int* kinds,
                   const uint64_t* as, const uint64_t* bs, const uint64_t* cs, uint64_t* out);

/* Deterministic generator used by the tests and the bench (SplitMix64). */
uint64_t so_splitmix64(uint64_t* state);
void so_fill_uniform(so_prime* const* primes, int nl, uint64_t seed,
                     uint64_t* x);

/* Timed CPU baseline helpers: run a relinearization of `count` ciphertexts. */
double so_now_seconds(void);

/* ---- wire formats and the ChaCha PRNG ---------------------------------- */
/* TranscribeBits (transcription.h:60-150) over integers held in uint64. */
long long so_transcribe_bits(const uint64_t* in, size_t in_len, int in_type_bits,
                             long long input_bit_length, int input_bits_per_int,
                             int output_bits_per_int, int out_type_bits, uint64_t* out,
                             size_t out_cap);
/* MontgomeryInt::SerializeVector / DeserializeVector (montgomery.cc:143-219). */
long long so_serialize_vector(const so_mont* p, const uint64_t* coeffs, size_t count,
                              uint8_t* out, size_t out_cap);
int so_deserialize_vector(const so_mont* p, int num_coeffs, const uint8_t* in, size_t in_len,
                          uint64_t* coeffs);
/* RFC 7539 ChaCha20 block; the reference's SingleThreadChaChaPrng
 * (prng/single_thread_chacha_prng.cc, prng/chacha_prng_util.cc). */
#define SO_CHACHA_OUTPUT_BYTES (255 * 32)
typedef struct so_chacha_prng {
  uint8_t key[32];
  int position;
  int salt_counter;
  uint8_t buffer[SO_CHACHA_OUTPUT_BYTES];
} so_chacha_prng;
void so_chacha20_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                       uint8_t out[64]);
int so_chacha_prng_init(so_chacha_prng* g, const uint8_t* key, size_t key_len);
uint8_t so_chacha_rand8(so_chacha_prng* g);
uint64_t so_chacha_rand64(so_chacha_prng* g);
uint64_t so_import_random_chacha(const so_mont* p, so_chacha_prng* g);
void so_sample_uniform_chacha(so_prime* const* primes, int nl, so_chacha_prng* g, uint64_t* x);

/* ---- BFV-side RNS operations (rns/rns_polynomial.cc:349-701) ---------- */
int so_poly_mod_reduce_msb(so_prime* const* primes, int nl, const uint64_t* ql_inv, int is_ntt,
                           uint64_t* x);
void so_switch_rns_basis(so_prime* const* in_p, int L, so_prime* const* out_p, int K,
                         const uint64_t* q_hat_inv, const uint64_t* q_hat_mod_ps,
                         const uint64_t* q_mod_ps, const uint64_t* in, uint64_t* out);
void so_scale_and_switch_rns_basis(so_prime* const* in_p, int L, so_prime* const* out_p, int K,
                                   const uint64_t* q_hat_inv, const uint64_t* q_inv_mod_ps,
                                   const uint64_t* p_mod_qs, const uint64_t* in, uint64_t* out);
void so_approx_switch_rns_basis(so_prime* const* in_p, int L, so_prime* const* out_p, int K,
                                const uint64_t* q_hat_inv, const uint64_t* q_hat_mod_ps,
                                int is_ntt, int balanced, const uint64_t* in, uint64_t* out);

/* ---- vectorised host baseline (shell_cpu_vec.c) ---- */
void so_hwy_ntt_fwd(const so_prime* pr, uint64_t* x);
void so_hwy_poly_ntt_fwd(so_prime* const* primes, int nl, uint64_t* x);
int so_hwy_relin_apply(so_prime* const* primes, int nl, const int* log_bs, int degree,
                       const uint64_t* comps, int comp_is_ntt_mask, const uint64_t* key_as,
                       const uint64_t* key_bs, uint64_t* out_c0, uint64_t* out_c1);

#ifdef __cplusplus
}
#endif
#endif
