/*
 * shell_gpu.h -- C-ABI of the MI355X (gfx950) implementation of SHELL's RNS
 * polynomial hot path: negacyclic NTT/INTT, MontgomeryInt arithmetic, BGV
 * modulus switching and gadget key switching (relinearization and Galois).
 *
 * The C++ API in shell-encryption_amd/cpp/rlwe/ (RnsPolynomial, RnsRelinKey,
 * RnsGaloisKey, RnsBgvCiphertext, ...) is a thin layer over these entry
 * points; other languages bind them directly (ctypes, cgo, JNI).
 *
 * Conventions
 *  - Every residue is a MontgomeryInt<Uint64> representation x*2^64 mod q,
 *    in [0, q) (or q for the image of NegateInPlace(0)), exactly as the
 *    reference stores it (shell_encryption/montgomery.h:226-666).
 *  - An RNS polynomial with L limbs and n = 2^log_n coefficients is L
 *    contiguous limbs of n uint64 words (limb-major), as the reference's
 *    coeff_vectors_ (shell_encryption/rns/rns_polynomial.h:627-635).  A batch
 *    of `count` polynomials is `count` such blocks back to back.
 *  - `primes` is a host array of L prime handles returned by
 *    shg_prime_create; limb i of every polynomial is reduced modulo primes[i].
 *    It plays the role of the reference's
 *    absl::Span<const PrimeModulus<ModularInt>* const> moduli.
 *  - Data pointers are device pointers (shg_alloc) unless named host_*.
 *  - All compute entry points are asynchronous on the context's stream (or
 *    the stream passed in); shg_sync waits for it.
 *  - Return value: SHG_OK (0) or a negative SHG_E* code; shg_last_error()
 *    returns the message, worded as the reference's absl::Status message
 *    where one exists.  On error nothing has been launched and outputs are
 *    unchanged (the reference's "leaves polynomial unchanged on failure").
 */
#ifndef SHELL_GPU_H_
#define SHELL_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHG_OK 0
#define SHG_E_INVALID_ARGUMENT -1   /* absl::InvalidArgumentError */
#define SHG_E_FAILED_PRECONDITION -2 /* absl::FailedPreconditionError */
#define SHG_E_INTERNAL -3           /* absl::InternalError */
#define SHG_E_DEVICE -4             /* HIP runtime error */
#define SHG_E_NO_MEMORY -5
#define SHG_E_UNSUPPORTED -6

#define SHG_MAX_LIMBS 64

typedef struct shg_context shg_context;

/* Pointwise operations (rns/rns_polynomial.h:159-365). */
typedef enum shg_pointwise_op {
  SHG_OP_ADD = 0, /* AddInPlace:  x <- x + a              */
  SHG_OP_SUB = 1, /* SubInPlace:  x <- x - a              */
  SHG_OP_MUL = 2, /* MulInPlace:  x <- x * a (REDC)       */
  SHG_OP_NEG = 3, /* NegateInPlace: x <- q - x (0 -> q)   */
  SHG_OP_FMA = 4  /* FusedMulAddInPlace: x <- x + a * b   */
} shg_pointwise_op;

/* NTT arithmetic variants, selectable per call (SHG_NTT_AUTO lets the
 * autotuner's table decide).  All variants are bit-exact. */
typedef enum shg_ntt_variant {
  SHG_NTT_AUTO = 0,
  SHG_NTT_INT_SHOUP = 1, /* 64-bit Shoup products from 32-bit mul_lo/mul_hi */
  SHG_NTT_F64 = 2        /* exact products from FP64 FMA (q < 2^50)        */
} shg_ntt_variant;
/* Launch-configuration flags that may be or-ed into a variant (the
 * autotuner reports its choice in this form): the alternate register tiling
 * of 2^12..2^14-word slices (8 words per thread below 2^14, 32 at 2^14,
 * instead of 16), and loading the next limb while transforming the current
 * one (FP64 only; ignored for the integer variant). */
#define SHG_NTT_TILING_ALT 4
#define SHG_NTT_PREFETCH 8
/* Forward transforms of 2^14-coefficient limbs split each limb over two
 * workgroups (both compute stage 0; each finishes one half), two of which
 * share a CU. */
#define SHG_NTT_SPLIT_LIMB 16
/* FP64 forward and inverse transforms of 2^12..2^14-coefficient limbs with
 * a single workgroup-wide exchange per limb; the other exchanges stay inside
 * each wave, so waves run independently between limbs.  Without the other
 * flags the FP64 inverse transform at 2^13 already is this one. */
#define SHG_NTT_WAVE_EXCHANGE 32
/* Stage split: 2^13..2^15-coefficient limbs as one radix-2 stage streamed
 * over HBM (first for the forward transform, last for the inverse) plus
 * 2^(log n - 1)-word slices, two per limb: twice the workgroups for batches
 * too small to fill the chip, at the cost of one more pass over the data. */
#define SHG_NTT_STAGE_SPLIT 64
/* FP64 forward and inverse transforms of 2^14-coefficient limbs with 32
 * words per thread and each LDS exchange in two halves, so that a
 * workgroup (one limb) needs 76 KiB of LDS and two of them share a CU: one
 * computes while the other loads or stores.  Without the other flags the
 * FP64 transforms at 2^14 already are these; other shapes and the integer
 * variant ignore the flag. */
#define SHG_NTT_TWO_PER_CU 128

/* ---- context, memory, streams ------------------------------------------ */
int shg_create(int device, shg_context** out);
int shg_destroy(shg_context* ctx);
const char* shg_last_error(void);
int shg_set_stream(shg_context* ctx, void* hip_stream); /* NULL = own stream */
void* shg_get_stream(shg_context* ctx);
int shg_sync(shg_context* ctx);
int shg_alloc(shg_context* ctx, size_t bytes, void** dptr);
int shg_free(shg_context* ctx, void* dptr);
int shg_upload(shg_context* ctx, void* dptr, const void* host, size_t bytes);
int shg_download(shg_context* ctx, void* host, const void* dptr, size_t bytes);
/* Asynchronous device-to-device copy on the context's stream. */
int shg_copy(shg_context* ctx, void* dst, const void* src, size_t bytes);

/* ---- primes: MontgomeryIntParams<Uint64>::Create (montgomery.cc:44-59) +
 *      InitializeNttParameters (ntt_parameters.h:183-235) ----------------- */
int shg_prime_create(shg_context* ctx, uint64_t q, int log_n, int* prime);
/* Copies the prime's host-side constants: params[10] = {modulus, log_modulus,
 * r_mod_modulus, r_mod_modulus_barrett, inv_modulus, inv_r, inv_r_barrett,
 * barrett_numerator, psi (Montgomery rep), n_inv (Montgomery rep)} and, if
 * tables != NULL, the 4*n words psis_bitrev_constant (.first, .second) then
 * psis_inv_bitrev_constant (.first, .second). */
int shg_prime_info(shg_context* ctx, int prime, uint64_t* params,
                   uint64_t* tables);

/* The ModReduceLsb constants of a chain: host_ql_inv[i] = Montgomery rep of
 * q_l^-1 mod q_i for i < num_limbs - 1, where q_l = primes[num_limbs - 1] is
 * the modulus being dropped (RnsContext::MainPrimeModulusInverseResidues,
 * rns/rns_context.h:145-148, built at rns/rns_context.cc:169-199).  Host
 * only; no device work. */
int shg_rns_ql_inv(shg_context* ctx, const int* primes, int num_limbs, uint64_t* host_ql_inv);

/* ---- transforms: RnsPolynomial::ConvertToNttForm / ConvertToCoeffForm
 *      (rns/rns_polynomial.cc:171-207), i.e. ForwardNumberTheoreticTransform
 *      and InverseNumberTheoreticTransform (dft_transformations.h:158-201)
 *      on every limb of `count` polynomials, in place. --------------------- */
int shg_ntt_fwd(shg_context* ctx, uint64_t* polys, int count, int num_limbs,
                const int* primes, int variant);
int shg_ntt_inv(shg_context* ctx, uint64_t* polys, int count, int num_limbs,
                const int* primes, int variant);

/* ---- pointwise: NegateInPlace/AddInPlace/SubInPlace/MulInPlace/
 *      FusedMulAddInPlace (rns/rns_polynomial.h:159-365) over `count`
 *      polynomials.  `a` / `b` are unused (may be NULL) where the op has no
 *      such operand.  a_stride / b_stride: words between consecutive
 *      polynomials of a and b (0 broadcasts one polynomial to all). ------- */
int shg_pointwise(shg_context* ctx, int op, uint64_t* x, const uint64_t* a,
                  size_t a_stride, const uint64_t* b, size_t b_stride,
                  int count, int num_limbs, const int* primes);
/* MulInPlace(Span<const ModularInt>) (rns/rns_polynomial.h:265-283):
 * host_scalars holds one Montgomery rep per limb. */
int shg_mul_scalar(shg_context* ctx, uint64_t* x, const uint64_t* host_scalars,
                   int count, int num_limbs, const int* primes);
/* MulInPlace(Int scalar) (rns/rns_polynomial.h:243-258): the integer is
 * imported modulo each limb's prime. */
int shg_mul_int(shg_context* ctx, uint64_t* x, uint64_t scalar, int count,
                int num_limbs, const int* primes);

/* ---- RnsPolynomial::Substitute (rns/rns_polynomial.cc:209-260): out may
 *      not alias in.  Polynomials must be in NTT form (checked by caller). */
int shg_substitute(shg_context* ctx, const uint64_t* in, uint64_t* out,
                   int power, int count, int num_limbs, const int* primes);

/* ---- RnsPolynomial::ModReduceLsb (rns/rns_polynomial.cc:262-344) in place:
 *      input polynomials have num_limbs limbs with stride num_limbs*n; on
 *      return limbs 0..num_limbs-2 of each hold the result (same stride).
 *      host_ql_inv: num_limbs-1 Montgomery reps of q_l^-1 mod q_i. -------- */
int shg_mod_reduce_lsb(shg_context* ctx, uint64_t* polys, uint64_t t,
                       const uint64_t* host_ql_inv, int is_ntt, int count,
                       int num_limbs, const int* primes);

/* ---- RnsPolynomial::ModReduceMsb (rns/rns_polynomial.cc:349-417) in place,
 *      the BFV modulus switch: limb i <- (a_i - lift([a]_{q_l})) * q_l^-1
 *      with the balanced lift of the dropped limb; layout and ql_inv as for
 *      shg_mod_reduce_lsb. */
int shg_mod_reduce_msb(shg_context* ctx, uint64_t* polys, const uint64_t* host_ql_inv,
                       int is_ntt, int count, int num_limbs, const int* primes);

/* ---- RNS basis conversion from the chain Q = prod q_i (num_in limbs) to
 *      P = prod p_j (num_out limbs), count polynomials, coefficient-form
 *      output; at most 16 moduli per chain.  All constants are Montgomery
 *      reps; host_q_hat_inv_mod_qs[i] = (Q/q_i)^-1 mod q_i.
 *  SHG_BASIS_EXACT:  SwitchRnsBasis (rns/rns_polynomial.cc:419-517), the
 *     centered representative of a mod P.  host_matrix[i*num_out + j] =
 *     (Q/q_i) mod p_j, host_vec[j] = Q mod p_j.
 *  SHG_BASIS_SCALE:  ScaleAndSwitchRnsBasis (rns/rns_polynomial.cc:519-636),
 *     round(P * a / Q) mod P.  host_matrix[j*num_in + i] = q_i^-1 mod p_j,
 *     host_vec[i] = P mod q_i.
 *  SHG_BASIS_APPROX: ApproxSwitchRnsBasis (rns/rns_polynomial.cc:638-701),
 *     sum_i [a_i (Q/q_i)^-1]_{q_i} (Q/q_i) mod p_j, the bracket lifted
 *     balanced when `balanced`; input may be in NTT form (in_is_ntt);
 *     host_matrix as for EXACT, host_vec unused.
 *  Coefficient-form input is required for EXACT and SCALE, as in the
 *  reference.  `out` must not overlap `in`. */
#define SHG_BASIS_EXACT 0
#define SHG_BASIS_SCALE 1
#define SHG_BASIS_APPROX 2
int shg_switch_rns_basis(shg_context* ctx, int mode, const uint64_t* in, int in_is_ntt,
                         int balanced, int count, int num_in, const int* in_primes,
                         uint64_t* out, int num_out, const int* out_primes,
                         const uint64_t* host_q_hat_inv_mod_qs, const uint64_t* host_matrix,
                         const uint64_t* host_vec);

/* ---- RnsRlweCiphertext::Mul (rns/rns_ciphertext.h:501-524): for each of
 *      `count` pairs, out (len_a+len_b-1 components) = tensor product of a
 *      (len_a components) and b (len_b components), all NTT form. -------- */
int shg_tensor_product(shg_context* ctx, const uint64_t* a, int len_a,
                       const uint64_t* b, int len_b, uint64_t* out, int count,
                       int num_limbs, const int* primes);

/* ---- key switching ------------------------------------------------------
 * RnsGadget (rns/rns_gadget.cc:42-153): host_log_bs[i] is log2 of the gadget
 * base for limb i; the gadget dimension is k = sum_i ceil(log q_i / log_bs[i]).
 * Keys are device arrays of k RNS polynomials (NTT form), key_as and key_bs,
 * as returned by RnsRelinKey::GetKeyA/GetKeyB and RnsGaloisKey::GetKeyA/
 * GetKeyB (rns/rns_relinearization_key.h:97-102, rns/rns_galois_key.h:214-219).
 */

/* RnsRelinKey::ApplyToRlweCiphertext (rns/rns_relinearization_key.cc:184-232)
 * for `count` ciphertexts of degree `degree` (degree+1 components each,
 * back to back).  Bit c of ntt_mask says component c is in NTT form (c0 and
 * c1 must be).  Key index (i-2)*degree + j as in the reference. */
int shg_relin_apply(shg_context* ctx, const uint64_t* comps, int degree,
                    int ntt_mask, const uint64_t* key_as,
                    const uint64_t* key_bs, int num_key_polys,
                    const int* host_log_bs, uint64_t* out_c0,
                    uint64_t* out_c1, int count, int num_limbs,
                    const int* primes);
/* shg_relin_apply without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Same arguments, checks and results as shg_relin_apply, for
 * ring degrees 2^4..2^14.  Its work space is a fixed 64 MiB area per GPU,
 * held from the library's load until the process ends; ciphertexts are
 * switched in chunks that fit it, and a call fails with
 * SHG_E_INVALID_ARGUMENT if one ciphertext alone needs more (gadget digit
 * planes of k*n*2 bytes, or L*n*8 bytes for bases above 2^16).  Every call
 * on one GPU, from any context or thread, shares that area: calls that
 * overlap on one GPU (different streams, or threads without ordering) race
 * and must be serialised by the caller. */
int shg_relin_apply_noalloc(shg_context* ctx, const uint64_t* comps, int degree,
                            int ntt_mask, const uint64_t* key_as,
                            const uint64_t* key_bs, int num_key_polys,
                            const int* host_log_bs, uint64_t* out_c0,
                            uint64_t* out_c1, int count, int num_limbs,
                            const int* primes);
/* RnsGaloisKey::ApplyToRlweCiphertext (rns/rns_galois_key.cc:238-287) for
 * `count` degree-1 ciphertexts (c0 then c1 per ciphertext). */
int shg_galois_apply(shg_context* ctx, const uint64_t* comps, int c1_is_ntt,
                     const uint64_t* key_as, const uint64_t* key_bs,
                     const int* host_log_bs, uint64_t* out_c0,
                     uint64_t* out_c1, int count, int num_limbs,
                     const int* primes);

/* A key-switching key prepared once for repeated use.  No counterpart in the
 * reference (its keys are RnsPolynomials kept by RnsRelinKey/RnsGaloisKey).
 * shg_ks_key_create keeps the caller's key_as/key_bs (num_key_polys
 * polynomials of num_limbs limbs; they must stay valid and unchanged while
 * the key is used) and, when every modulus is below 2^50, device copies of
 * them with each residue held as an FP64 value (exact), which the FP64
 * digit-plane key switch at n = 2^13 and 2^14 multiplies without converting
 * (two operations fewer per key word); other shapes read the caller's
 * words.  The results equal shg_relin_apply's / shg_galois_apply's word for
 * word.  Creation synchronises the context's stream and refuses an active
 * capture; the apply calls may be captured like the plain ones. */
typedef struct shg_ks_key shg_ks_key;
int shg_ks_key_create(shg_context* ctx, const uint64_t* key_as,
                      const uint64_t* key_bs, int num_key_polys,
                      int num_limbs, const int* primes, shg_ks_key** out);
int shg_ks_key_destroy(shg_context* ctx, shg_ks_key* key);
/* shg_relin_apply / shg_galois_apply with a prepared key (its moduli must be
 * `primes`). */
int shg_relin_apply_key(shg_context* ctx, const uint64_t* comps, int degree,
                        int ntt_mask, const shg_ks_key* key,
                        const int* host_log_bs, uint64_t* out_c0,
                        uint64_t* out_c1, int count, int num_limbs,
                        const int* primes);
int shg_galois_apply_key(shg_context* ctx, const uint64_t* comps, int c1_is_ntt,
                         const shg_ks_key* key, const int* host_log_bs,
                         uint64_t* out_c0, uint64_t* out_c1, int count,
                         int num_limbs, const int* primes);

/* RnsGadget::Decompose (rns/rns_gadget.cc:108-153): `a` (coefficient form)
 * -> k digit polynomials (coefficient form), count inputs. */
int shg_gadget_decompose(shg_context* ctx, const uint64_t* a,
                         const int* host_log_bs, uint64_t* digits, int count,
                         int num_limbs, const int* primes);

/* ---- LazyRnsPolynomial: num_terms FusedMulAddInPlace then Export
 *      (rns/lazy_rns_polynomial.h:110-129, rns/lazy_rns_polynomial.cc:134-151):
 *      out = sum_t as[t] * bs[t] (Montgomery product), NTT form.  as_stride /
 *      bs_stride are words between consecutive terms. ------------------- */
int shg_lazy_fma(shg_context* ctx, const uint64_t* as, size_t as_stride,
                 const uint64_t* bs, size_t bs_stride, int num_terms,
                 uint64_t* out, int num_limbs, const int* primes);

/* LazyRnsPolynomial::FusedMulSumAddInPlace / FusedMulDifferenceAddInPlace
 * (rns/lazy_rns_polynomial.cc:170-254): x <- x + (a + b) * c, or with
 * subtract != 0 x <- x + (a - b) * c (Montgomery products, NTT form), count
 * polynomials of num_limbs limbs laid out back to back.  x stays canonical
 * after every call, which leaves the residue the reference's 128-bit lazy
 * sum exports.  Moduli of 2^60 and above are refused with the reference's
 * message (:73-78). */
int shg_fused_mul_sum_add(shg_context* ctx, int subtract, uint64_t* x, const uint64_t* a,
                          const uint64_t* b, const uint64_t* c, int count, int num_limbs,
                          const int* primes);

/* ---- wire formats --------------------------------------------------------
 * RnsPolynomial::Serialize / Deserialize (rns/rns_polynomial.h:99-140) store
 * one byte string per limb, MontgomeryInt::SerializeVector /
 * DeserializeVector (montgomery.cc:143-219): the n representations of limb i
 * as consecutive log2(q_i)-bit fields of a little-endian bit stream
 * (TranscribeBits, transcription.h:60-150), ceil(n * log2(q_i) / 8) bytes.
 * Here one polynomial's strings lie back to back; *poly_bytes is their sum
 * and limb_bytes (may be NULL) receives each string's length.  n <= 2^15
 * (kMaxNumCoeffs). */
int shg_serialized_bytes(shg_context* ctx, const int* primes, int num_limbs,
                         size_t* limb_bytes, size_t* poly_bytes);
/* count polynomials (device words) -> count * poly_bytes device bytes. */
int shg_serialize_polys(shg_context* ctx, const uint64_t* polys, int count,
                        int num_limbs, const int* primes, uint8_t* out);
/* The inverse, straight into device buffers; like the reference, the fields
 * are taken as representations without a range check. */
int shg_deserialize_polys(shg_context* ctx, const uint8_t* in, int count,
                          int num_limbs, const int* primes, uint64_t* polys);
/* RnsPolynomial::SampleUniform (rns/rns_polynomial.h:640-667) of `count`
 * polynomials drawn in order from one SingleThreadChaChaPrng seeded with the
 * 32-byte host_seed (prng/single_thread_chacha_prng.cc:36-62): the random
 * pad that RnsGaloisKey regenerates from its seed
 * (rns/rns_galois_key.cc:81-105, 204-236).  Output is limb-major per
 * polynomial (NTT form, as the reference marks it); moduli of 9..62 bits. */
int shg_sample_uniform_chacha(shg_context* ctx, const uint8_t* host_seed, size_t seed_len,
                              int count, int num_limbs, const int* primes, uint64_t* out);

/* ---- encrypted queries against a device-resident database --------------
 * Oblivious expansion (ObliviousExpander::ExpandCiphertext,
 * oblivious_expand.h:173-227, there on the non-RNS stack) on RNS
 * ciphertexts: host_cts[g] points to query g's buffer of 2^levels
 * ciphertexts (each c0 then c1, NTT form, num_limbs limbs of n words);
 * ciphertext 0 holds the query.  Level j (0-based) turns ciphertexts
 * 0..2^j-1 into 0..2^(j+1)-1: sub = RnsGaloisKey::ApplyTo(Substitute(ct,
 * n/2^j + 1)), ct[k + 2^j] = (ct[k] - sub) * X^(-2^j), ct[k] += sub, with
 * the Galois key for power n/2^j + 1 (key_as / key_bs: k polynomials each,
 * NTT form).  All queries of a call share the level. */
int shg_oblivious_expand_level(shg_context* ctx, uint64_t* const* host_cts, int count,
                               int level, const uint64_t* key_as, const uint64_t* key_bs,
                               const int* host_log_bs, int num_limbs, const int* primes);
/* shg_oblivious_expand_level with the level's Galois key prepared
 * (shg_ks_key_create); same results. */
int shg_oblivious_expand_level_key(shg_context* ctx, uint64_t* const* host_cts,
                                   int count, int level, const shg_ks_key* key,
                                   const int* host_log_bs, int num_limbs,
                                   const int* primes);
/* All levels 0..levels-1; key_as / key_bs are host arrays of one device
 * key pointer per level. */
int shg_oblivious_expand(shg_context* ctx, uint64_t* const* host_cts, int count, int levels,
                         const uint64_t* const* key_as, const uint64_t* const* key_bs,
                         const int* host_log_bs, int num_limbs, const int* primes);
/* Database records are RNS plaintext polynomials in NTT form (num_limbs
 * limbs of n words each).  The inner product reads them exported
 * (ExportInt of every word), which shg_pir_db_export writes. */
int shg_pir_db_export(shg_context* ctx, const uint64_t* db, long long num_records,
                      uint64_t* out, int num_limbs, const int* primes);
/* Synthetic records for benchmarks: coefficients uniform in [0, t) from a
 * seeded SplitMix64 stream, imported mod each limb and transformed to NTT
 * form. */
int shg_pir_db_synthesize(shg_context* ctx, uint64_t seed, uint64_t t, long long num_records,
                          uint64_t* out, int num_limbs, const int* primes);
/* For each query g: answer_g[c] = sum_{r < 2^levels} expanded_g[r] (.)
 * record[r * num_columns + c] (AbsorbInPlace then AddInPlace,
 * rns/rns_ciphertext.h:219-225), c < num_columns; host_out[g] receives
 * num_columns ciphertexts.  Moduli below 2^50. */
int shg_pir_inner_product(shg_context* ctx, const uint64_t* const* host_expanded, int count,
                          int levels, const uint64_t* db_exported, int num_columns,
                          uint64_t* const* host_out, int num_limbs, const int* primes);

/* ---- serving: continuous batching of encrypted queries ------------------
 * No counterpart in the reference.  A server owns a worker thread that
 * admits queued requests while a slot is free and their HBM footprint
 * ((2^levels + max columns) ciphertexts each) fits the budget, then
 * advances every admitted request by one stage per step: all requests at
 * expansion level j share one shg_oblivious_expand_level call, all fully
 * expanded requests of one database one shg_pir_inner_product call.
 * Requests arriving meanwhile join at the next step.  Keys (one Galois key
 * per level, power n/2^j + 1) and databases (exported records, see
 * shg_pir_db_export; row r column c is record r * num_columns + c) stay
 * resident and are not owned by the server.  A request's answer
 * (num_columns ciphertexts) depends only on its query. */
typedef struct shg_pir_server shg_pir_server;
typedef struct shg_pir_server_config {
  int log_n;               /* ring degree of queries and records */
  int levels;              /* expansion levels (1..log_n): 2^levels database rows */
  int max_active;          /* slots: most requests on the device at once */
  size_t hbm_budget_bytes; /* per-request buffer budget; 0 = slots only */
  int cohort;              /* while requests run, admit newcomers only when this
                              many can start together (0 or 1: at once) */
} shg_pir_server_config;
int shg_pir_server_create(shg_context* ctx, const int* primes, int num_limbs,
                          const int* host_log_bs, const uint64_t* const* key_as,
                          const uint64_t* const* key_bs, const shg_pir_server_config* cfg,
                          shg_pir_server** out);
int shg_pir_server_add_db(shg_pir_server* s, const uint64_t* db_exported, int num_columns,
                          int* db_id);
/* Starts the worker; databases cannot be added afterwards. */
int shg_pir_server_start(shg_pir_server* s);
/* Queues a query (host copy of one ciphertext, c0 then c1, NTT form). */
int shg_pir_server_submit(shg_pir_server* s, int db_id, const uint64_t* host_query,
                          long long* request);
/* Blocks until the request is answered; host_answer (may be NULL) receives
 * num_columns ciphertexts, latency_ms the time from submission. */
int shg_pir_server_wait(shg_pir_server* s, long long request, uint64_t* host_answer,
                        double* latency_ms);
int shg_pir_server_stats(shg_pir_server* s, long long* steps, long long* stage_runs,
                         long long* batched_items, int* capacity);
/* Submits num_requests queries (cycling over num_distinct host queries) at
 * Poisson arrivals of rate_per_s from a thread of its own, waits for all and
 * returns each latency and the wall time. */
int shg_pir_server_load_test(shg_pir_server* s, int db_id, const uint64_t* host_queries,
                             int num_distinct, int num_requests, double rate_per_s,
                             uint64_t seed, double* latencies_ms, double* wall_s);
/* Stops the worker (queued and running requests fail) and frees the slots.
 * No other thread may be waiting on the server. */
int shg_pir_server_destroy(shg_pir_server* s);

/* ---- batched uneven work (persistent launch) ---------------------------
 * No counterpart in the reference, which transforms one RnsPolynomial limb
 * at a time (rns/rns_polynomial.cc:171-207): these run the per-limb
 * operations of many polynomials of different ring degree and level in one
 * launch.  One job = one limb of n = 2^log_n words at `data` (device,
 * 16-byte aligned), 2^11 <= n <= 2^15, reduced modulo `prime` (whose log_n
 * must match).  Jobs must not overlap. */
typedef struct shg_limb_job {
  uint64_t* data;
  const uint64_t* other; /* second operand for multiply jobs, else NULL */
  int prime;
  int log_n;
} shg_limb_job;
#define SHG_BATCH_NTT_FWD 0 /* ForwardNumberTheoreticTransform */
#define SHG_BATCH_NTT_INV 1 /* InverseNumberTheoreticTransform */
#define SHG_BATCH_MUL 2     /* MulInPlace: data <- data * other (REDC) */
typedef struct shg_batch_plan shg_batch_plan;
/* Validates the jobs, uploads the job table (sorted by decreasing size) and
 * sizes the persistent grid.  variant as for shg_ntt_fwd (AUTO takes the
 * autotuner's choice for the plan's largest ring). */
int shg_batch_plan_create(shg_context* ctx, const shg_limb_job* host_jobs,
                          int num_jobs, int variant, shg_batch_plan** out);
int shg_batch_plan_destroy(shg_context* ctx, shg_batch_plan* plan);
/* One polynomial of a batched modulus switch: RnsPolynomial::ModReduceLsb
 * (rns/rns_polynomial.cc:262-344) in place, as shg_mod_reduce_lsb does for
 * one shape; num_limbs <= 16. */
typedef struct shg_poly_job {
  uint64_t* data;              /* num_limbs limbs of n words, 16-byte aligned */
  const int* primes;           /* host array of num_limbs prime handles */
  const uint64_t* host_ql_inv; /* num_limbs - 1 Montgomery reps of q_l^-1 mod q_i */
  int num_limbs;
  int log_n;
} shg_poly_job;
#define SHG_BATCH_MOD_REDUCE 3
int shg_batch_mod_reduce_plan_create(shg_context* ctx, const shg_poly_job* host_jobs,
                                     int num_jobs, uint64_t t, int is_ntt, int variant,
                                     shg_batch_plan** out);
/* One kernel launch on the context's stream, nothing else: may be captured
 * in a graph (shg_graph_begin).  A limb plan takes SHG_BATCH_NTT_FWD,
 * SHG_BATCH_NTT_INV or SHG_BATCH_MUL, a modulus-switch plan
 * SHG_BATCH_MOD_REDUCE. */
int shg_batch_run(shg_context* ctx, const shg_batch_plan* plan, int op);
/* Plan, run and wait in one call. */
int shg_batch_limbs(shg_context* ctx, const shg_limb_job* host_jobs,
                    int num_jobs, int op);

/* ---- HIP graphs of op chains --------------------------------------------
 * Captures every launch the context makes between begin and end into a
 * graph (stream capture on the context's stream); launch replays it.
 * May be captured: shg_copy, the transforms, pointwise ops, scalar
 * multiplies, substitute, tensor product, lazy FMA, gadget decomposition,
 * shg_batch_run, shg_mod_reduce_lsb, shg_relin_apply and shg_galois_apply.
 * The last three use work buffers held by the context: make the same call
 * once outside capture at the largest size the graph will use, so that no
 * buffer has to grow; a call that would have to grow one during capture
 * fails with SHG_E_FAILED_PRECONDITION.  While any graph of the context is
 * alive, a buffer that an eager call grows is retired rather than freed, so
 * replays never touch freed memory; retired buffers are released when the
 * last graph is destroyed.  Entry points that synchronise with the host,
 * allocate, or copy from host memory (shg_sync, shg_alloc, shg_free,
 * shg_upload, shg_download, shg_prime_create, plan creation and
 * destruction, shg_batch_limbs, oblivious expansion, the inner product,
 * profiling queries, the autotuner, shg_graph_destroy) fail with
 * SHG_E_FAILED_PRECONDITION while the stream is being captured. */
typedef struct shg_graph shg_graph;
int shg_graph_begin(shg_context* ctx);
int shg_graph_end(shg_context* ctx, shg_graph** out);
int shg_graph_launch(shg_context* ctx, const shg_graph* graph);
int shg_graph_destroy(shg_context* ctx, shg_graph* graph);

/* ---- timing -------------------------------------------------------------
 * Per-kernel timing with HIP events recorded on the context's stream around
 * each launch (off by default; the reference has no tracing of its own).
 * Names: "ntt_fwd", "ntt_inv", "pointwise", "substitute", "tensor_product",
 * "keyswitch_intt", "keyswitch". */
int shg_profile_enable(shg_context* ctx, int enable);
int shg_profile_reset(shg_context* ctx);
int shg_profile_query(shg_context* ctx, const char* kernel, double* total_ms,
                      long long* launches);

/* ---- autotuner ----------------------------------------------------------
 * No counterpart in the reference.  Times the NTT launch configurations
 * that apply to a shape (forward + inverse pairs, best of three interleaved
 * rounds after a warm-up) and records the fastest for SHG_NTT_AUTO calls:
 *   arithmetic      FP64 (moduli below 2^50) or 64-bit integer Shoup;
 *   kernel family   register/LDS passes, wave exchange (SHG_NTT_WAVE_EXCHANGE),
 *                   split forward limbs (SHG_NTT_SPLIT_LIMB, 2^14);
 *   register radix  16 words per thread or the alternate 8 / 32
 *                   (SHG_NTT_TILING_ALT, 2^12..2^14 slices);
 *   stage split     whole-limb slices or one streamed stage plus half-limb
 *                   slices (SHG_NTT_STAGE_SPLIT, 2^13..2^15);
 *   prefetch        loading the next limb during the current (FP64).
 * The table is keyed by (log n, limbs per polynomial, total limbs rounded
 * down to a power of two, modulus width: below 2^50 or not); a call looks up
 * its own key, then the same key for any limbs per polynomial.
 * shg_autotune_ntt tunes `limbs` limbs on one synthetic modulus below 2^50
 * (one limb per polynomial; recorded for every limb count of that total, and
 * for batch plans of that ring degree).  *chosen receives the configuration
 * code: the variant in bits 0-1 or-ed with the flags above. */
int shg_autotune_ntt(shg_context* ctx, int log_n, int limbs, int* chosen);
/* As shg_autotune_ntt for `count` polynomials over the caller's moduli
 * (their width selects the candidates; uniform residues of each modulus). */
int shg_autotune_ntt_moduli(shg_context* ctx, const int* primes, int num_limbs, int count,
                            int* chosen);

#ifdef __cplusplus
}
#endif
#endif /* SHELL_GPU_H_ */
