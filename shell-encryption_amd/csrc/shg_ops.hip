// Elementwise, permutation, modulus-switch, gadget and lazy-accumulation
// kernels for gfx950.  All are HBM-streaming kernels: one thread per 64-bit
// residue, consecutive lanes on consecutive words, grid sized to cover the
// batch (thousands of workgroups for the configured batches).
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"

namespace shg {

static inline unsigned grid_for(long long items, int block) {
  return (unsigned)((items + block - 1) / block);
}

// ---------------------------------------------------------------------------
// n^-1 scaling (the tail of InverseNumberTheoreticTransform,
// dft_transformations.h:195-199) for limbs whose transform ran in pieces.
// ---------------------------------------------------------------------------
__global__ void scale_ninv_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                                  LimbPrimes lp, int num_limbs, int logn, int mode,
                                  long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long row = i >> logn;
  const PrimeDev& P = primes[lp.id[row % num_limbs]];
  const uint64_t c = mode == 2 ? P.n_inv_x_c : P.n_inv_c;
  const uint64_t cb = mode == 2 ? P.n_inv_x_cb : P.n_inv_cb;
  data[i] = csub(shoup_lazy(data[i], c, cb, P.q), P.q);
}

hipError_t launch_scale_ninv(uint64_t* data, const PrimeDev* primes, const LimbPrimes& lp, int rows,
                             int num_limbs, int logn, int mode, hipStream_t s) {
  const long long total = (long long)rows << logn;
  hipLaunchKernelGGL(scale_ninv_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, data, primes,
                     lp, num_limbs, logn, mode, total);
  return hipGetLastError();
}

// Zero fill by vector stores: a kernel node rather than a memset node when the
// launch is captured into a graph.
__global__ void zero_kernel(uint64_t* __restrict__ x, long long words) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) x[i] = 0;
}

hipError_t launch_zero(uint64_t* x, size_t words, hipStream_t s) {
  if (words == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for((long long)words, 256)), dim3(256), 0, s, x,
                     (long long)words);
  return hipGetLastError();
}

__global__ void export_kernel(uint64_t* __restrict__ x, const PrimeDev* __restrict__ primes,
                              LimbPrimes lp, int num_limbs, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int limb = (int)((i >> logn) % num_limbs);
  x[i] = mont_export(x[i], primes[lp.id[limb]]);
}

hipError_t launch_export(uint64_t* x, const PrimeDev* primes, const LimbPrimes& lp, int count,
                         int num_limbs, int logn, hipStream_t s) {
  const long long total = ((long long)count * num_limbs) << logn;
  hipLaunchKernelGGL(export_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, x, primes, lp,
                     num_limbs, logn, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pointwise ops (rns/rns_polynomial.h:159-365 over montgomery.h:318-465).
// ---------------------------------------------------------------------------
template <int OP>
__global__ void pointwise_kernel(uint64_t* __restrict__ x, const uint64_t* __restrict__ a,
                                 size_t a_stride, const uint64_t* __restrict__ b, size_t b_stride,
                                 const PrimeDev* __restrict__ primes, LimbPrimes lp, int num_limbs,
                                 int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly_words = (long long)num_limbs << logn;
  const long long poly = i / poly_words;
  const long long off = i - poly * poly_words;
  const int limb = (int)(off >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  const uint64_t q = P.q;
  uint64_t v = x[i];
  if (OP == SHG_OP_NEG) {
    v = q - v;
  } else {
    const uint64_t av = a[poly * a_stride + off];
    if (OP == SHG_OP_ADD) v = mod_add(v, av, q);
    if (OP == SHG_OP_SUB) v = mod_sub(v, av, q);
    if (OP == SHG_OP_MUL) v = mont_mul(v, av, P);
    if (OP == SHG_OP_FMA) v = mont_fma(v, av, b[poly * b_stride + off], P);
  }
  x[i] = v;
}

hipError_t launch_pointwise(int op, uint64_t* x, const uint64_t* a, size_t a_stride,
                            const uint64_t* b, size_t b_stride, const PrimeDev* primes,
                            const LimbPrimes& lp, int count, int num_limbs, int logn,
                            hipStream_t s) {
  const long long total = ((long long)count * num_limbs) << logn;
  const dim3 g(grid_for(total, 256)), blk(256);
  switch (op) {
    case SHG_OP_ADD:
      hipLaunchKernelGGL(pointwise_kernel<SHG_OP_ADD>, g, blk, 0, s, x, a, a_stride, b, b_stride,
                         primes, lp, num_limbs, logn, total);
      break;
    case SHG_OP_SUB:
      hipLaunchKernelGGL(pointwise_kernel<SHG_OP_SUB>, g, blk, 0, s, x, a, a_stride, b, b_stride,
                         primes, lp, num_limbs, logn, total);
      break;
    case SHG_OP_MUL:
      hipLaunchKernelGGL(pointwise_kernel<SHG_OP_MUL>, g, blk, 0, s, x, a, a_stride, b, b_stride,
                         primes, lp, num_limbs, logn, total);
      break;
    case SHG_OP_NEG:
      hipLaunchKernelGGL(pointwise_kernel<SHG_OP_NEG>, g, blk, 0, s, x, a, a_stride, b, b_stride,
                         primes, lp, num_limbs, logn, total);
      break;
    case SHG_OP_FMA:
      hipLaunchKernelGGL(pointwise_kernel<SHG_OP_FMA>, g, blk, 0, s, x, a, a_stride, b, b_stride,
                         primes, lp, num_limbs, logn, total);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// x <- x + (a + b) * c or x + (a - b) * c (Montgomery product), the
// per-term update of LazyRnsPolynomial::FusedMulSumAddInPlace /
// FusedMulDifferenceAddInPlace (rns/lazy_rns_polynomial.cc:170-254).  The
// reference adds the integer products (a + b) c and (a + q - b) c into a
// 128-bit sum that Export reduces and takes out of Montgomery form; keeping
// x reduced after every term gives the same residue, canonical.
template <bool SUB>
__global__ void fused_mul_sum_add_kernel(uint64_t* __restrict__ x, const uint64_t* __restrict__ a,
                                         const uint64_t* __restrict__ b,
                                         const uint64_t* __restrict__ c,
                                         const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                         int num_limbs, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int limb = (int)((i >> logn) % num_limbs);
  const PrimeDev& P = primes[lp.id[limb]];
  const uint64_t s = SUB ? mod_sub(a[i], b[i], P.q) : mod_add(a[i], b[i], P.q);
  x[i] = mont_fma(x[i], s, c[i], P);
}

hipError_t launch_fused_mul_sum_add(bool subtract, uint64_t* x, const uint64_t* a,
                                    const uint64_t* b, const uint64_t* c, const PrimeDev* primes,
                                    const LimbPrimes& lp, int count, int num_limbs, int logn,
                                    hipStream_t s) {
  const long long total = ((long long)count * num_limbs) << logn;
  const dim3 g(grid_for(total, 256)), blk(256);
  if (subtract)
    hipLaunchKernelGGL(fused_mul_sum_add_kernel<true>, g, blk, 0, s, x, a, b, c, primes, lp,
                       num_limbs, logn, total);
  else
    hipLaunchKernelGGL(fused_mul_sum_add_kernel<false>, g, blk, 0, s, x, a, b, c, primes, lp,
                       num_limbs, logn, total);
  return hipGetLastError();
}

__global__ void key_to_f64_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                  const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                  int num_limbs, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint64_t q = primes[lp.id[(i >> logn) % num_limbs]].q;
  const uint64_t w = in[i];
  const double v = w > (q >> 1) ? u52_to_f64(w) - u52_to_f64(q) : u52_to_f64(w);
  out[i] = (uint64_t)__double_as_longlong(v);
}

hipError_t launch_key_to_f64(const uint64_t* in, uint64_t* out, const PrimeDev* primes,
                             const LimbPrimes& lp, int count, int num_limbs, int logn,
                             hipStream_t s) {
  const long long total = ((long long)count * num_limbs) << logn;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(key_to_f64_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, in, out,
                     primes, lp, num_limbs, logn, total);
  return hipGetLastError();
}

__global__ void mul_scalar_kernel(uint64_t* __restrict__ x, LimbPrimes slo, LimbPrimes shi,
                                  const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                  int num_limbs, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int limb = (int)((i >> logn) % num_limbs);
  const PrimeDev& P = primes[lp.id[limb]];
  const uint64_t sc = ((uint64_t)(uint32_t)shi.id[limb] << 32) | (uint32_t)slo.id[limb];
  x[i] = mont_mul(x[i], sc, P);
}

hipError_t launch_mul_scalar(uint64_t* x, const LimbPrimes& slo, const LimbPrimes& shi,
                             const PrimeDev* primes, const LimbPrimes& lp, int count, int num_limbs,
                             int logn, hipStream_t s) {
  const long long total = ((long long)count * num_limbs) << logn;
  hipLaunchKernelGGL(mul_scalar_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, x, slo, shi,
                     primes, lp, num_limbs, logn, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Substitute (rns/rns_polynomial.cc:209-260): out[brv(j)] = in[brv(idx_j)]
// with idx_j = ((power-1)/2 + j*power) mod n.  Computed as a gather: each
// output word finds its source directly, so the permutation is one pass.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned brev_n(unsigned x, int logn) {
  return __brev(x) >> (32 - logn);
}

__global__ void substitute_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                  int power, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const unsigned n = 1u << logn;
  const long long base = i & ~(long long)(n - 1);
  const unsigned pos = (unsigned)(i & (n - 1));  // output position = brv(j)
  const unsigned j = brev_n(pos, logn);
  const unsigned idx = (unsigned)((((unsigned long long)(power - 1) / 2) +
                                   (unsigned long long)j * (unsigned)power) & (n - 1));
  out[i] = in[base + brev_n(idx, logn)];
}

hipError_t launch_substitute(const uint64_t* in, uint64_t* out, int power, int count,
                             int num_limbs, int logn, hipStream_t s) {
  const long long total = ((long long)count * num_limbs) << logn;
  hipLaunchKernelGGL(substitute_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, in, out,
                     power, logn, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ModReduceLsb (rns/rns_polynomial.cc:262-344).
// ---------------------------------------------------------------------------
__global__ void modreduce_extract_kernel(const uint64_t* __restrict__ polys,
                                         uint64_t* __restrict__ delta_l,
                                         const PrimeDev* __restrict__ primes, int last_prime,
                                         uint64_t k_mod_ql, int num_limbs, int logn,
                                         long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = i >> logn;
  const long long off = i & ((1LL << logn) - 1);
  const PrimeDev& P = primes[last_prime];
  const uint64_t v = polys[((poly * num_limbs + (num_limbs - 1)) << logn) + off];
  delta_l[i] = mont_mul(v, k_mod_ql, P);
}

hipError_t launch_modreduce_extract(const uint64_t* polys, uint64_t* delta_l,
                                    const PrimeDev* primes, int last_prime, uint64_t k_mod_ql,
                                    int count, int num_limbs, int logn, hipStream_t s) {
  const long long total = (long long)count << logn;
  hipLaunchKernelGGL(modreduce_extract_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, polys,
                     delta_l, primes, last_prime, k_mod_ql, num_limbs, logn, total);
  return hipGetLastError();
}

__global__ void modreduce_lift_kernel(const uint64_t* __restrict__ delta_l,
                                      uint64_t* __restrict__ delta,
                                      const PrimeDev* __restrict__ primes, int last_prime,
                                      LimbPrimes lp, int nr, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long row = i >> logn;  // poly * nr + limb
  const long long poly = row / nr;
  const int limb = (int)(row - poly * nr);
  const long long off = i & ((1LL << logn) - 1);
  delta[i] = convert_balanced(delta_l[(poly << logn) + off], primes[last_prime],
                              primes[lp.id[limb]]);
}

hipError_t launch_modreduce_lift(const uint64_t* delta_l, uint64_t* delta, const PrimeDev* primes,
                                 int last_prime, const LimbPrimes& lp, int count, int num_limbs,
                                 int logn, hipStream_t s) {
  const int nr = num_limbs - 1;
  const long long total = ((long long)count * nr) << logn;
  hipLaunchKernelGGL(modreduce_lift_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, delta_l,
                     delta, primes, last_prime, lp, nr, logn, total);
  return hipGetLastError();
}

__global__ void modreduce_combine_kernel(uint64_t* __restrict__ polys,
                                         const uint64_t* __restrict__ delta, ModReduceArgs args,
                                         const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                         int num_limbs, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int nr = num_limbs - 1;
  const long long row = i >> logn;
  const long long poly = row / nr;
  const int limb = (int)(row - poly * nr);
  const long long off = i & ((1LL << logn) - 1);
  const PrimeDev& P = primes[lp.id[limb]];
  uint64_t* xp = polys + (((poly * num_limbs) + limb) << logn) + off;
  const uint64_t a = mont_mul(*xp, args.r_mod_qi[limb], P);       // r * a
  const uint64_t d = mont_mul(delta[i], args.t_mod_qi[limb], P);  // t * delta
  *xp = mont_mul(mod_add(a, d, P.q), args.ql_inv[limb], P);
}

hipError_t launch_modreduce_combine(uint64_t* polys, const uint64_t* delta,
                                    const ModReduceArgs& args, const PrimeDev* primes,
                                    const LimbPrimes& lp, int count, int num_limbs, int logn,
                                    hipStream_t s) {
  const int nr = num_limbs - 1;
  const long long total = ((long long)count * nr) << logn;
  hipLaunchKernelGGL(modreduce_combine_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, polys,
                     delta, args, primes, lp, num_limbs, logn, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tensor product (rns/rns_ciphertext.h:501-524): out[m] = sum_{i+j=m} a_i*b_j,
// every output word computed once from its len_a*len_b operand words.
// ---------------------------------------------------------------------------
__global__ void tensor_kernel(const uint64_t* __restrict__ a, int len_a,
                              const uint64_t* __restrict__ b, int len_b, uint64_t* __restrict__ out,
                              const PrimeDev* __restrict__ primes, LimbPrimes lp, int num_limbs,
                              int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly_words = (long long)num_limbs << logn;
  const long long ct = i / poly_words;
  const long long off = i - ct * poly_words;
  const int limb = (int)(off >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  const uint64_t* ap = a + ct * len_a * poly_words + off;
  const uint64_t* bp = b + ct * len_b * poly_words + off;
  uint64_t av[8], bv[8];
  for (int u = 0; u < len_a; ++u) av[u] = ap[u * poly_words];
  for (int v = 0; v < len_b; ++v) bv[v] = bp[v * poly_words];
  uint64_t* op = out + ct * (len_a + len_b - 1) * poly_words + off;
  for (int m = 0; m < len_a + len_b - 1; ++m) {
    uint64_t acc = 0;
    for (int u = 0; u < len_a; ++u) {
      const int v = m - u;
      if (v >= 0 && v < len_b) acc = mont_fma(acc, av[u], bv[v], P);
    }
    op[m * poly_words] = acc;
  }
}

hipError_t launch_tensor_product(const uint64_t* a, int len_a, const uint64_t* b, int len_b,
                                 uint64_t* out, const PrimeDev* primes, const LimbPrimes& lp,
                                 int count, int num_limbs, int logn, hipStream_t s) {
  if (len_a > 8 || len_b > 8) return hipErrorInvalidValue;
  const long long total = ((long long)count * num_limbs) << logn;
  hipLaunchKernelGGL(tensor_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, a, len_a, b,
                     len_b, out, primes, lp, num_limbs, logn, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gadget decomposition (rns/rns_gadget.cc:108-153 with gadget.h:52-79 and the
// balanced lift of modulus_conversion.h:30-83).  A digit of limb i is below
// min(2^log_b, q_i); one above q_i / 2 (possible when log_b >= log q_i - 1)
// stands for digit - q_i on the other limbs.
// ---------------------------------------------------------------------------
__global__ void gadget_decompose_kernel(const uint64_t* __restrict__ a,
                                        uint64_t* __restrict__ digits, GadgetDesc gd,
                                        const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                        int num_limbs, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  // i indexes (ct, digit j, target limb l, coefficient).
  const long long row = i >> logn;
  const long long off = i & ((1LL << logn) - 1);
  const int l = (int)(row % num_limbs);
  const long long rj = row / num_limbs;
  const int j = (int)(rj % gd.k);
  const long long ct = rj / gd.k;
  const int src = gd.src_limb[j];
  const PrimeDev& Ps = primes[lp.id[src]];
  const uint64_t v = mont_export(a[((ct * num_limbs + src) << logn) + off], Ps);
  const uint64_t digit = (v >> gd.shift[j]) & ((1ull << gd.log_b[j]) - 1);
  digits[i] = import_balanced(digit, Ps.q, primes[lp.id[l]]);
}

hipError_t launch_gadget_decompose(const uint64_t* a, uint64_t* digits, const GadgetDesc& gd,
                                   const PrimeDev* primes, const LimbPrimes& lp, int count,
                                   int num_limbs, int logn, hipStream_t s) {
  const long long total = ((long long)count * gd.k * num_limbs) << logn;
  hipLaunchKernelGGL(gadget_decompose_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, a,
                     digits, gd, primes, lp, num_limbs, logn, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LazyRnsPolynomial (rns/lazy_rns_polynomial.h:84-129,
// rns/lazy_rns_polynomial.cc:134-151): 128-bit accumulation of raw products,
// Barrett-reduced only where the reference would refresh, then exported.
// ---------------------------------------------------------------------------
// (hi * 2^64 + lo) mod q: ImportInt(hi) is exactly hi * 2^64 mod q.
__device__ __forceinline__ uint64_t reduce_u128(uint64_t lo, uint64_t hi, const PrimeDev& P) {
  return mod_add(mont_import(hi, P), reduce_full(lo, P.q, P.bn), P.q);
}

__global__ void lazy_fma_kernel(const uint64_t* __restrict__ as, size_t as_stride,
                                const uint64_t* __restrict__ bs, size_t bs_stride, int num_terms,
                                uint64_t* __restrict__ out, const PrimeDev* __restrict__ primes,
                                LimbPrimes lp, int logn, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int limb = (int)(i >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  uint64_t lo = 0, hi = 0;
  // Products of residues below 2^log_q are below 2^(2 log_q), so
  // 2^(127 - 2 log_q) of them fit in 128 bits (the reference's maximal level,
  // rns/lazy_rns_polynomial.h:156-181); fold the accumulator back below q
  // at that interval.  Folding leaves the residue unchanged.
  int lvl = 127 - 2 * P.log_q;
  lvl = lvl > 30 ? 30 : (lvl < 1 ? 1 : lvl);
  const int refresh_mask = (1 << lvl) - 1;
  for (int t = 0; t < num_terms; ++t) {
    const uint64_t a = as[(size_t)t * as_stride + i];
    const uint64_t b = bs[(size_t)t * bs_stride + i];
    const uint64_t plo = a * b;
    const uint64_t phi = mulhi64(a, b);
    lo += plo;
    hi += phi + (lo < plo ? 1 : 0);
    if ((t & refresh_mask) == refresh_mask) {
      lo = reduce_u128(lo, hi, P);
      hi = 0;
    }
  }
  out[i] = mont_export(reduce_u128(lo, hi, P), P);
}

hipError_t launch_lazy_fma(const uint64_t* as, size_t as_stride, const uint64_t* bs,
                           size_t bs_stride, int num_terms, uint64_t* out, const PrimeDev* primes,
                           const LimbPrimes& lp, int num_limbs, int logn, hipStream_t s) {
  const long long total = (long long)num_limbs << logn;
  hipLaunchKernelGGL(lazy_fma_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, as, as_stride,
                     bs, bs_stride, num_terms, out, primes, lp, logn, total);
  return hipGetLastError();
}

}  // namespace shg
