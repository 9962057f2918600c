// Host-callable launchers for the gfx950 kernels (internal to libshellgpu).
#ifndef SHG_KERNELS_H_
#define SHG_KERNELS_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shg_device.h"

namespace shg {

// Largest LDS slice (log2 words) a transform workgroup holds: 2^14 words
// (136 KiB with padding) fits the 160 KiB of LDS per CU.
constexpr int kMaxLogM = 14;

// inv_mode (inverse only): 1 = times n^-1 (InverseNumberTheoreticTransform),
// 2 = times n^-1 * 2^-64 (the inverse transform followed by ExportInt).
// NTT launch configuration: the arithmetic variant in bits 0-1
// (SHG_NTT_INT_SHOUP or SHG_NTT_F64), the alternate register tiling of
// 2^12..2^14-word slices in kNttCfgAlt, next-item prefetch (FP64 only) in
// kNttCfgPrefetch, forward transforms of 2^14-word limbs split over two
// workgroups in kNttCfgSplit, FP64 forward transforms of 2^12..2^14-word
// limbs with one workgroup-wide exchange (the rest inside each wave) in
// kNttCfgWave, one streamed radix-2 stage plus half-limb slices for
// 2^13..2^15-word limbs in kNttCfgHalf, FP64 transforms of 2^14-word limbs
// with 32 words per lane and two workgroups per CU in kNttCfgPair.  A bare
// variant is a valid configuration.
constexpr int kNttCfgAlt = 4;
constexpr int kNttCfgPrefetch = 8;
constexpr int kNttCfgSplit = 16;
constexpr int kNttCfgWave = 32;
constexpr int kNttCfgHalf = 64;
constexpr int kNttCfgPair = 128;
hipError_t launch_ntt(bool inverse, int cfg, uint64_t* data, const PrimeDev* primes,
                      const LimbPrimes& lp, int rows, int num_limbs, int logn, int max_logm,
                      hipStream_t s, int inv_mode = 1);
hipError_t launch_scale_ninv(uint64_t* data, const PrimeDev* primes, const LimbPrimes& lp,
                             int rows, int num_limbs, int logn, int mode, hipStream_t s);
// x[0 .. words) <- 0 with a kernel (never a memset node in a captured graph).
hipError_t launch_zero(uint64_t* x, size_t words, hipStream_t s);
// x <- ExportInt(x) (montgomery.h:156-164) over count polynomials.
hipError_t launch_export(uint64_t* x, const PrimeDev* primes, const LimbPrimes& lp, int count,
                         int num_limbs, int logn, hipStream_t s);

hipError_t launch_pointwise(int op, uint64_t* x, const uint64_t* a, size_t a_stride,
                            const uint64_t* b, size_t b_stride, const PrimeDev* primes,
                            const LimbPrimes& lp, int count, int num_limbs, int logn,
                            hipStream_t s);
hipError_t launch_fused_mul_sum_add(bool subtract, uint64_t* x, const uint64_t* a,
                                    const uint64_t* b, const uint64_t* c, const PrimeDev* primes,
                                    const LimbPrimes& lp, int count, int num_limbs, int logn,
                                    hipStream_t s);
// x[limb] *= scalars[limb] (Montgomery reps; REDC product).
hipError_t launch_mul_scalar(uint64_t* x, const LimbPrimes& scalars_lo,
                             const LimbPrimes& scalars_hi, const PrimeDev* primes,
                             const LimbPrimes& lp, int count, int num_limbs, int logn,
                             hipStream_t s);
hipError_t launch_substitute(const uint64_t* in, uint64_t* out, int power, int count,
                             int num_limbs, int logn, hipStream_t s);

// ModReduceLsb building blocks.
struct ModReduceArgs {
  uint64_t t;
  uint64_t k_mod_ql;                 // ImportInt(k) mod q_l (Montgomery rep)
  uint64_t r_mod_qi[SHG_MAX_LIMBS];  // ImportInt(r) mod q_i
  uint64_t t_mod_qi[SHG_MAX_LIMBS];  // ImportInt(t) mod q_i
  uint64_t ql_inv[SHG_MAX_LIMBS];    // q_l^-1 mod q_i (Montgomery reps)
};
// delta_l (scratch, count limbs) <- last limb of each polynomial times k.
hipError_t launch_modreduce_extract(const uint64_t* polys, uint64_t* delta_l,
                                    const PrimeDev* primes, int last_prime, uint64_t k_mod_ql,
                                    int count, int num_limbs, int logn, hipStream_t s);
// delta (count x (num_limbs-1) limbs) <- balanced conversion of delta_l.
hipError_t launch_modreduce_lift(const uint64_t* delta_l, uint64_t* delta,
                                 const PrimeDev* primes, int last_prime, const LimbPrimes& lp,
                                 int count, int num_limbs, int logn, hipStream_t s);
// polys[i] <- (r * polys[i] + t * delta[i]) * ql_inv[i] for i < num_limbs-1.
hipError_t launch_modreduce_combine(uint64_t* polys, const uint64_t* delta,
                                    const ModReduceArgs& args, const PrimeDev* primes,
                                    const LimbPrimes& lp, int count, int num_limbs, int logn,
                                    hipStream_t s);

hipError_t launch_tensor_product(const uint64_t* a, int len_a, const uint64_t* b, int len_b,
                                 uint64_t* out, const PrimeDev* primes, const LimbPrimes& lp,
                                 int count, int num_limbs, int logn, hipStream_t s);

// Gadget description: digit j comes from limb src_limb[j], bit offset
// shift[j], mask (1<<log_b[j])-1.
struct GadgetDesc {
  int k;
  unsigned char src_limb[256];
  unsigned char shift[256];
  unsigned char log_b[256];
  // Per source limb: digit count and base log (digits of limb i are
  // consecutive in j).
  unsigned char limb_digits[SHG_MAX_LIMBS];
  unsigned char limb_log_b[SHG_MAX_LIMBS];
  unsigned char limb_first[SHG_MAX_LIMBS];  // global index of limb i's first digit
  int max_log_b;
  // 2^max_log_b * 64 <= every modulus: digits enter the transform far below
  // q, which lets the FP64 key switch defer its first reduction.
  int small_digits;
  // Per source limb: a digit can exceed q_i / 2 (base 2^b with
  // 2^b - 1 > q_i / 2), so its lift to the other limbs must take the
  // balanced rule of ConvertModulusBalanced (modulus_conversion.h:30-83):
  // d > q_i / 2 stands for d - q_i.
  unsigned char limb_balance[SHG_MAX_LIMBS];
  int any_balance;
  // Digit bits of packed planes at n = 2^14 (plane_pack_bits): the common
  // base log when every limb has the same one and it is 8, 10 or 12; else 0.
  int plane_bits;
};
hipError_t launch_gadget_decompose(const uint64_t* a, uint64_t* digits, const GadgetDesc& gd,
                                   const PrimeDev* primes, const LimbPrimes& lp, int count,
                                   int num_limbs, int logn, hipStream_t s);

// Fused key switch: for each (ciphertext, output limb), for every gadget
// digit, extract + forward NTT in LDS and multiply-accumulate with the key;
// acc0 starts from c0 (or zero), acc1 from c1 (or zero).
struct KeySwitchArgs {
  const uint64_t* src;       // exported (ExportInt) coefficient-form input, count x L x n
  size_t src_stride;         // words between ciphertexts in src
  const uint16_t* planes;    // or: digit planes (launch_ntt_inv_digits), count x k x n
  const uint64_t* c0;        // NTT-form c0 (added), or null
  size_t c0_stride;
  const uint64_t* c1;        // NTT-form c1 (added), or null
  size_t c1_stride;
  const uint64_t* key_a;     // key polys, index key_base + j
  const uint64_t* key_b;
  int key_base;
  uint64_t* out0;
  uint64_t* out1;
  size_t out_stride;
  int accumulate;            // 1: out += ..., else out = init + ...
  int key_f64;               // 1: key words are FP64 values (launch_key_to_f64)
};
// out = the residues of `count` polynomials (num_limbs limbs of 2^logn words)
// as FP64 values in balanced form (w - q for w > q/2; exact below 2^50): the
// key format the FP64 key switch reads without conversion.
hipError_t launch_key_to_f64(const uint64_t* in, uint64_t* out, const PrimeDev* primes,
                             const LimbPrimes& lp, int count, int num_limbs, int logn,
                             hipStream_t s);
hipError_t launch_keyswitch(const KeySwitchArgs& ka, const GadgetDesc& gd, int variant,
                            const PrimeDev* primes, const LimbPrimes& lp, int count,
                            int num_limbs, int logn, hipStream_t s);

hipError_t launch_lazy_fma(const uint64_t* as, size_t as_stride, const uint64_t* bs,
                           size_t bs_stride, int num_terms, uint64_t* out,
                           const PrimeDev* primes, const LimbPrimes& lp, int num_limbs,
                           int logn, hipStream_t s);

// Records msg as the calling thread's last error (shg_last_error) and
// returns code.
int set_error(int code, const char* msg);

// Gadget digits of the inverse transform of a component, written as 16-bit
// planes for the key switch: limb l of ciphertext c is read at
// in + c * in_stride + l * n (NTT form), transformed to coefficient form
// with ExportInt folded into the n^-1 scaling, and digit dd of limb l goes
// to plane (c * k + limb_first[l] + dd), which stores position p at index
// (p % TC) * VC + p / TC with VC = 16 (n <= 2^13) or 2^(SHG_KS_LOGE+1)
// (n = 2^14) and
// TC = n / VC: a key-switch thread reads its VC digits contiguously.
// Requires every limb's base log <= 16 and 4 <= log n <= 14.
hipError_t launch_ntt_inv_digits(int variant, const uint64_t* in, size_t in_stride,
                                 uint16_t* planes, const GadgetDesc& gd, const PrimeDev* primes,
                                 const LimbPrimes& lp, int count, int num_limbs, int logn,
                                 hipStream_t s);
// Words per thread (log2) of the key switch at n = 2^14 (two 2^13-word
// slices): 8 words with 1024 threads (four waves per SIMD, four exchanges
// per digit) measured 8.73 against 8.82 ms per 512 ciphertexts at L = 8
// for 16 words with 512 threads (two waves per SIMD, three exchanges), and
// its digit planes (16 entries per thread) store faster: 0.27 vs 0.32 ms.
#ifndef SHG_KS_LOGE
#define SHG_KS_LOGE 3
#endif
constexpr int ks_planes_vc(int logn) { return logn >= 14 ? (2 << SHG_KS_LOGE) : 16; }
// FP64 digit planes at n = 2^14 read by keyswitch_n14_kernel: row r (16
// entries, one key-switch thread's digits) holds limb positions
// (r >> 6) + 16 (r & 63) + 1024 k, k = 0..15; the row of position p is
// plane_row14(p & 1023).  Without it row r holds r + 1024 k.
#ifndef SHG_KS_N14
#define SHG_KS_N14 1
#endif
__host__ __device__ constexpr int plane_row14(int p) {
  return SHG_KS_N14 ? 64 * (p & 15) + ((p >> 4) & 63) : p;
}
// Packed planes at n = 2^14: with every digit of b = gd.plane_bits bits a
// row's 16 digits take W = ceil(16 b / 32) dwords (digit i at bits
// [b i, b i + b)), and dword k of row r sits at dword k * 1024 + r of the
// plane's slot (the slot keeps its 2^14 16-bit entries; the rest is unused).
#ifndef SHG_KS_PACK
#define SHG_KS_PACK 1
#endif
__host__ __device__ constexpr int plane_pack_bits(int plane_bits) {
  return SHG_KS_PACK && SHG_KS_N14 ? plane_bits : 0;
}

// Sets hipFuncAttributeMaxDynamicSharedMemorySize once per kernel and device.
void set_max_dynamic_shmem(const void* kernel, size_t bytes);

// Persistent launch over uneven limb jobs (2^11 <= n <= 2^15).
struct LimbJobDev {
  uint64_t* data;
  const uint64_t* other;
  int prime;
  int log_n;
};
constexpr int kBatchThreads = 1024;
constexpr int kBatchMinLogN = 11;
constexpr int kBatchMaxLogN = 15;
constexpr int kBatchNttFwd = 0;
constexpr int kBatchNttInv = 1;
constexpr int kBatchMul = 2;
constexpr int kBatchModReduce = 3;
constexpr int kBatchMaxPolyLimbs = 16;
// One polynomial of the batched modulus switch (ModReduceArgs per job).
struct PolyJobDev {
  uint64_t* data;
  int log_n;
  int num_limbs;
  int is_ntt;
  int prime[kBatchMaxPolyLimbs];
  uint64_t k_mod_ql;
  uint64_t r_mod_qi[kBatchMaxPolyLimbs];
  uint64_t t_mod_qi[kBatchMaxPolyLimbs];
  uint64_t ql_inv[kBatchMaxPolyLimbs];
  // FP64 form of the same constants (q < 2^50), plain residues: the
  // inverse transform of the last limb is scaled by k * n^-1 * 2^-64 (giving
  // the exported k * a_l directly), and limb i becomes
  // a_i * c1[i] + NTT(lift_i(delta)) * c2[i] with c1 = r * q_l^-1 * 2^-128
  // and c2 = t * q_l^-1 * 2^-64 (Montgomery reps r, t, q_l^-1).
  double f_kl;
  double f_c1[kBatchMaxPolyLimbs];
  double f_c2[kBatchMaxPolyLimbs];
};
size_t batch_shmem_bytes(int max_log_n);
const void* batch_kernel_ptr(int variant, bool modreduce);
// scratch: 2^16 words per workgroup of the grid.
hipError_t launch_batch_modreduce(const PolyJobDev* jobs, int num_jobs, int variant,
                                  const PrimeDev* primes, uint64_t* scratch, int grid,
                                  size_t shmem, hipStream_t s);
// ModReduceLsb of `count` equal polynomials (job `uni`, data at uni.data +
// j * stride) with the fused per-polynomial kernel: the grid from
// modreduce_uniform_grid (which also sets the kernel's LDS limit), scratch
// of 2^16 words per workgroup.
int modreduce_uniform_grid(int variant, int log_n, int count);
hipError_t launch_modreduce_uniform(const PolyJobDev& uni, size_t stride, int count, int variant,
                                    const PrimeDev* primes, uint64_t* scratch, int grid,
                                    hipStream_t s);
// The same for FP64 at n = 2^14 in NTT form on the two-workgroups-per-CU
// transforms: two launches, scratch of count * 2^14 words.
hipError_t launch_modreduce_w32(const PolyJobDev& uni, size_t stride, int count,
                                const PrimeDev* primes, uint64_t* scratch, hipStream_t s);
hipError_t launch_batch_limbs(const LimbJobDev* jobs, int num_jobs, int op, int variant,
                              const PrimeDev* primes, int grid, size_t shmem, hipStream_t s);
// FP64 forward / inverse (x n^-1) transforms of `rows` limbs of 2^15 words,
// each limb read and written once (shg_batch.hip).
hipError_t launch_ntt15(bool inverse, uint64_t* data, const PrimeDev* primes,
                        const LimbPrimes& lp, int rows, int num_limbs, hipStream_t s);

// Wire formats (shg_wire.hip).  A polynomial's serialized form is its limbs'
// byte strings back to back: limb i holds n residues of bits[i] = log2(q_i)
// bits each, little-endian bit stream, at byte limb_off[i] of the
// polynomial; limb_off[num_limbs] is the polynomial's size.  `aligned`: every
// limb string is a multiple of 8 bytes long and the buffer is 8-byte
// aligned (64-bit loads and stores).
struct WireDesc {
  int num_limbs;
  int log_n;
  int aligned;
  unsigned char bits[SHG_MAX_LIMBS];
  unsigned long long limb_off[SHG_MAX_LIMBS + 1];
};
hipError_t launch_pack(const uint64_t* polys, uint8_t* out, const WireDesc& wd, int count,
                       hipStream_t s);
hipError_t launch_unpack(const uint8_t* in, uint64_t* polys, const WireDesc& wd, int count,
                         hipStream_t s);
// RnsPolynomial::SampleUniform of `count` polynomials from one
// SingleThreadChaChaPrng stream (key = the 32-byte seed as 8 little-endian
// words); every modulus has 9..62 bits.
struct PadDesc {
  uint32_t key[8];
  int count;
  int num_limbs;
  int log_n;
  unsigned char bits[SHG_MAX_LIMBS];
  uint64_t q[SHG_MAX_LIMBS];
};
hipError_t launch_chacha_pad(const PadDesc& pd, uint64_t* out, hipStream_t s);

// RNS basis conversions (shg_bfv.hip): modes as SHG_BASIS_* of shell_gpu.h.
constexpr int kBasisMax = 16;
constexpr int kBasisExact = 0;
constexpr int kBasisScale = 1;
constexpr int kBasisApprox = 2;
struct BasisArgs {
  int mode;
  int nin, nout;
  int balanced;  // approximate conversion with the balanced lift
  int in_id[kBasisMax], out_id[kBasisMax];
  uint64_t q_hat_inv[kBasisMax];        // (Q/q_i)^-1 mod q_i
  uint64_t pre[kBasisMax];              // scale: P mod q_i
  uint64_t mat[kBasisMax * kBasisMax];  // [i * nout + j]: (Q/q_i) mod p_j, or -(q_i^-1) mod p_j
  uint64_t vec[kBasisMax];              // exact: Q mod p_j
};
// Coefficient-form input (count polynomials of nin limbs) -> count
// polynomials of nout limbs, coefficient form.
hipError_t launch_basis_switch(const uint64_t* in, uint64_t* out, const BasisArgs& ba,
                               const PrimeDev* primes, int count, int logn, hipStream_t s);

// Oblivious expansion and the inner product against a database (shg_pir.hip).
hipError_t launch_expand_substitute(const uint64_t* const* cts, int count, int m, uint64_t* S,
                                    int power, int logn, int num_limbs, hipStream_t s);
hipError_t launch_expand_update(uint64_t* const* cts, int count, int m, const uint64_t* sub0,
                                const uint64_t* sub1, const uint64_t* mono,
                                const PrimeDev* primes, const LimbPrimes& lp, int logn,
                                int num_limbs, hipStream_t s);
int pir_k_split(int count, int C, int K, int logn, int num_limbs);
// expd: device array of the count expanded-query pointers; expd0: the first
// of them as known on the host (the single-query kernel takes it directly).
hipError_t launch_pir_inner_product(const uint64_t* const* expd, const uint64_t* expd0,
                                    int count, int K, const uint64_t* db_x, int C,
                                    uint64_t* part, int KS,
                                    uint64_t* const* out, const PrimeDev* primes,
                                    const LimbPrimes& lp, int logn, int num_limbs, hipStream_t s);
hipError_t launch_export_words(const uint64_t* in, uint64_t* out, long long polys,
                               const PrimeDev* primes, const LimbPrimes& lp, int logn,
                               int num_limbs, hipStream_t s);
hipError_t launch_db_synth(uint64_t* out, long long records, uint64_t seed, uint64_t t,
                           const PrimeDev* primes, const LimbPrimes& lp, int logn, int num_limbs,
                           hipStream_t s);

}  // namespace shg

#endif  // SHG_KERNELS_H_
