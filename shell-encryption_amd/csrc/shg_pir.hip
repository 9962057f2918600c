// Oblivious expansion and the plaintext-by-ciphertext inner product of an
// encrypted query against a device-resident database (RNS ciphertexts).
//
// Expansion follows ObliviousExpander::ExpandCiphertext
// (shell_encryption/oblivious_expand.h:173-227) on the RNS stack: at level
// j, with m = 2^j ciphertexts per query and substitution power
// p_j = n / 2^j + 1,
//   sub          = GaloisKeySwitch(Substitute(result[k], p_j))
//   result[k+m]  = (result[k] - sub) * X^(-2^j)
//   result[k]    = result[k] + sub
// using RnsPolynomial::Substitute (rns/rns_polynomial.cc:209-260),
// RnsGaloisKey::ApplyTo (rns/rns_galois_key.cc:238-287), SubInPlace,
// AddInPlace and AbsorbInPlace (a MulInPlace by the NTT-form monomial,
// rns/rns_ciphertext.h:219-225).  Queries are processed together: every
// kernel runs over all (query, ciphertext) pairs of a level, the key switch
// over one packed array.
//
// The inner product answer[c] = sum_r expanded[r] (.) record[r * C + c]
// (AbsorbInPlace then AddInPlace per term) is computed with exact FP64
// products of the expanded residues and the records' exported residues
// (x * 2^-64 mod q): sum_r e_r * (d_r 2^-64) is the Montgomery product sum
// the reference forms, reduced once per eight terms.
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"

namespace shg {

namespace {

unsigned grid_of(long long total, int block) { return (unsigned)((total + block - 1) / block); }

__device__ __forceinline__ unsigned brev_bits(unsigned x, int bits) {
  return __brev(x) >> (32 - bits);
}

// S[(g*m + k)*2 + comp] = Substitute(result_g[k][comp], power) for every
// query g, active ciphertext k < m and component; packed.
__global__ void expand_substitute_kernel(const uint64_t* const* __restrict__ cts, int m,
                                         uint64_t* __restrict__ S, int power, int logn,
                                         int num_limbs, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = (long long)num_limbs << logn;
  const long long ct = i / (2 * poly);  // g * m + k
  const long long rem = i - ct * 2 * poly;
  const int g = (int)(ct / m);
  const int k = (int)(ct - (long long)g * m);
  const unsigned n = 1u << logn;
  const long long limb_base = rem & ~(long long)(n - 1);
  const unsigned pos = (unsigned)(rem & (n - 1));
  const unsigned j = brev_bits(pos, logn);
  const unsigned idx = (unsigned)((((unsigned long long)(power - 1) / 2) +
                                   (unsigned long long)j * (unsigned)power) & (n - 1));
  const uint64_t* src = cts[g] + (long long)k * 2 * poly + limb_base;
  S[i] = src[brev_bits(idx, logn)];
}

// result_g[k + m] = (result_g[k] - sub) * mono, result_g[k] += sub, both
// components, for every query g and k < m.  sub0 / sub1 are the packed key
// switch outputs (index g*m + k).
__global__ void expand_update_kernel(uint64_t* const* __restrict__ cts, int m,
                                     const uint64_t* __restrict__ sub0,
                                     const uint64_t* __restrict__ sub1,
                                     const uint64_t* __restrict__ mono,
                                     const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                     int logn, int num_limbs, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = (long long)num_limbs << logn;
  const long long ct = i / poly;  // g * m + k
  const long long off = i - ct * poly;
  const int g = (int)(ct / m);
  const int k = (int)(ct - (long long)g * m);
  const int limb = (int)(off >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  uint64_t* r = cts[g] + (long long)k * 2 * poly + off;
  uint64_t* w = cts[g] + (long long)(k + m) * 2 * poly + off;
  const uint64_t s0 = sub0[i], s1 = sub1[i];
  const uint64_t r0 = r[0], r1 = r[poly];
  const uint64_t mo = mono[off];
  w[0] = mont_mul(mod_sub(r0, s0, P.q), mo, P);
  w[poly] = mont_mul(mod_sub(r1, s1, P.q), mo, P);
  r[0] = mod_add(r0, s0, P.q);
  r[poly] = mod_add(r1, s1, P.q);
}

// Queries per thread for a batch of `count` (each record word loaded feeds
// 2 * QG products) and columns per thread (each expanded word feeds that
// many): 1 x 4 for a single query (its own kernel), 2 x 8 and 4 x 4 (16
// accumulator pairs) for batches.  Measured at 2^18 records x 64 columns: 4.5 ms for one
// query; 3.1, 2.5 and 2.3 ms per query for 2, 4 and 16 queries (4.1-4.2 ms
// per query with one query per thread).
int pir_query_group(int count) { return count >= 4 ? 4 : (count >= 2 ? 2 : 1); }
constexpr int pir_columns_per_thread(int qg) { return qg == 2 ? 8 : 4; }

// Partial inner products of one query: thread (column block of four, K-chunk,
// limb, pos), offsets fastest.  part[((g * C + c) * KS + ks) * 2 + comp]
// [limb][pos] holds the canonical sum over the chunk's terms.
__global__ void pir_partial_single_kernel(const uint64_t* __restrict__ ex, int K,
                                          const uint64_t* __restrict__ db_x, int C, int KS,
                                          uint64_t* __restrict__ part,
                                          const PrimeDev* __restrict__ primes, LimbPrimes lp,
                                          int logn, int num_limbs, long long total) {
  constexpr int kCB = 4;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = (long long)num_limbs << logn;
  const long long off = i % poly;
  const long long rest = i / poly;
  const int ks = (int)(rest % KS);
  const int cb = (int)(rest / KS);
  const int limb = (int)(off >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  const double q = P.qd, qinv = P.qinv;
  const int chunk = (K + KS - 1) / KS;
  const int r0 = ks * chunk;
  const int r1 = r0 + chunk < K ? r0 + chunk : K;
  const int c0 = cb * kCB;
  double a0[kCB], a1[kCB];
#pragma unroll
  for (int cc = 0; cc < kCB; ++cc) a0[cc] = a1[cc] = 0.0;
  int since = 0;
  for (int r = r0; r < r1; ++r) {
    const double e0 = u52_to_f64(ex[(long long)r * 2 * poly + off]);
    const double e1 = u52_to_f64(ex[(long long)r * 2 * poly + poly + off]);
#pragma unroll
    for (int cc = 0; cc < kCB; ++cc) {
      if (c0 + cc < C) {
        const double w = u52_to_f64(db_x[((long long)r * C + c0 + cc) * poly + off]);
        const double wq = w * qinv;
        a0[cc] += f64_mulmod(e0, w, wq, q);
        a1[cc] += f64_mulmod(e1, w, wq, q);
      }
    }
    // Terms are below 0.875 q (canonical operands, q < 2^50): eight of them
    // on a reduced accumulator stay below 7.5 q < 2^53.
    if (++since == 8) {
      since = 0;
#pragma unroll
      for (int cc = 0; cc < kCB; ++cc) {
        a0[cc] = f64_reduce(a0[cc], q, qinv);
        a1[cc] = f64_reduce(a1[cc], q, qinv);
      }
    }
  }
#pragma unroll
  for (int cc = 0; cc < kCB; ++cc) {
    if (c0 + cc >= C) continue;
    const long long base = ((long long)(c0 + cc) * KS + ks) * 2 * poly + off;
    part[base] = f64_canon_u52(f64_reduce(a0[cc], q, qinv), q);
    part[base + poly] = f64_canon_u52(f64_reduce(a1[cc], q, qinv), q);
  }
}

// Partial inner products of a batch: a thread accumulates QG queries x kCB
// columns x 2 components for one word offset (limb * n + pos) and one
// K-chunk of rows, so each record word it loads feeds 2 * QG products and
// each expanded word kCB.  A workgroup owns a tile of kTile consecutive
// offsets for one (query group, column block, chunk); tiles are dealt out so
// that all workgroups of a tile have indices congruent mod 8 and consecutive
// otherwise: under round-robin dispatch they run together on one XCD, and
// the records and expanded rows they share come from that XCD's L2 rather
// than once per workgroup from HBM.  part as above, per query g.
constexpr int kTile = 256;

template <int QG, int kCB = pir_columns_per_thread(QG)>
__global__ void __launch_bounds__(kTile)
    pir_partial_kernel(const uint64_t* const* __restrict__ expd, int count, int K,
                       const uint64_t* __restrict__ db_x, int C, int KS,
                       uint64_t* __restrict__ part, const PrimeDev* __restrict__ primes,
                       LimbPrimes lp, int logn, int num_limbs) {
  const long long poly = (long long)num_limbs << logn;
  const int cb_count = (C + kCB - 1) / kCB;
  const int tiles = (int)((poly + kTile - 1) / kTile);
  const int per_tile = ((count + QG - 1) / QG) * cb_count * KS;
  const int b = blockIdx.x;
  const int j = b >> 3;
  const int tile = (b & 7) + 8 * (j / per_tile);
  if (tile >= tiles) return;
  int rest = j % per_tile;
  const int ks = rest % KS;
  rest /= KS;
  const int cb = rest % cb_count;
  const int g0 = (rest / cb_count) * QG;
  const long long off = (long long)tile * kTile + threadIdx.x;
  if (off >= poly) return;
  const int nq = count - g0 < QG ? count - g0 : QG;
  const int limb = (int)(off >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  const double q = P.qd, qinv = P.qinv;
  const int chunk = (K + KS - 1) / KS;
  const int r0 = ks * chunk;
  const int r1 = r0 + chunk < K ? r0 + chunk : K;
  const int c0 = cb * kCB;
  double a0[QG][kCB], a1[QG][kCB];
#pragma unroll
  for (int jq = 0; jq < QG; ++jq) {
#pragma unroll
    for (int cc = 0; cc < kCB; ++cc) a0[jq][cc] = a1[jq][cc] = 0.0;
  }
  const uint64_t* ex[QG];
#pragma unroll
  for (int jq = 0; jq < QG; ++jq) ex[jq] = jq < nq ? expd[g0 + jq] : expd[g0];
  int since = 0;
  for (int r = r0; r < r1; ++r) {
    double w[kCB], wq[kCB];
#pragma unroll
    for (int cc = 0; cc < kCB; ++cc) {
      w[cc] = c0 + cc < C ? u52_to_f64(db_x[((long long)r * C + c0 + cc) * poly + off]) : 0.0;
      wq[cc] = w[cc] * qinv;
    }
#pragma unroll
    for (int jq = 0; jq < QG; ++jq) {
      if (jq < nq) {
        const double e0 = u52_to_f64(ex[jq][(long long)r * 2 * poly + off]);
        const double e1 = u52_to_f64(ex[jq][(long long)r * 2 * poly + poly + off]);
#pragma unroll
        for (int cc = 0; cc < kCB; ++cc) {
          a0[jq][cc] += f64_mulmod(e0, w[cc], wq[cc], q);
          a1[jq][cc] += f64_mulmod(e1, w[cc], wq[cc], q);
        }
      }
    }
    // Terms are below 0.875 q (canonical operands, q < 2^50): eight of them
    // on a reduced accumulator stay below 7.5 q < 2^53.
    if (++since == 8) {
      since = 0;
#pragma unroll
      for (int jq = 0; jq < QG; ++jq) {
#pragma unroll
        for (int cc = 0; cc < kCB; ++cc) {
          a0[jq][cc] = f64_reduce(a0[jq][cc], q, qinv);
          a1[jq][cc] = f64_reduce(a1[jq][cc], q, qinv);
        }
      }
    }
  }
#pragma unroll
  for (int jq = 0; jq < QG; ++jq) {
    if (jq >= nq) continue;
#pragma unroll
    for (int cc = 0; cc < kCB; ++cc) {
      if (c0 + cc >= C) continue;
      const long long base = (((long long)(g0 + jq) * C + c0 + cc) * KS + ks) * 2 * poly + off;
      part[base] = f64_canon_u52(f64_reduce(a0[jq][cc], q, qinv), q);
      part[base + poly] = f64_canon_u52(f64_reduce(a1[jq][cc], q, qinv), q);
    }
  }
}

// answer_g[c][comp] = sum over KS partials, canonical.
__global__ void pir_sum_kernel(const uint64_t* __restrict__ part, int C, int KS,
                               uint64_t* const* __restrict__ out,
                               const PrimeDev* __restrict__ primes, LimbPrimes lp, int logn,
                               int num_limbs, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = (long long)num_limbs << logn;
  const long long off = i % (2 * poly);  // comp * poly + limb * n + pos
  const long long gc = i / (2 * poly);   // g * C + c
  const int g = (int)(gc / C);
  const int c = (int)(gc - (long long)g * C);
  const int limb = (int)((off % poly) >> logn);
  const PrimeDev& P = primes[lp.id[limb]];
  uint64_t s = 0;
  for (int ks = 0; ks < KS; ++ks) s = mod_add(s, part[(gc * KS + ks) * 2 * poly + off], P.q);
  out[g][(long long)c * 2 * poly + off] = s;
}

__global__ void export_words_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                    const PrimeDev* __restrict__ primes, LimbPrimes lp, int logn,
                                    int num_limbs, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = (long long)num_limbs << logn;
  const int limb = (int)((i % poly) >> logn);
  out[i] = mont_export(in[i], primes[lp.id[limb]]);
}

// Records with coefficients uniform in [0, t) from SplitMix64 of (seed,
// record, coefficient), imported mod each limb (coefficient form).
__global__ void db_synth_kernel(uint64_t* __restrict__ out, uint64_t seed, uint64_t t,
                                const PrimeDev* __restrict__ primes, LimbPrimes lp, int logn,
                                int num_limbs, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long poly = (long long)num_limbs << logn;
  const long long rec = i / poly;
  const long long off = i - rec * poly;
  const int limb = (int)(off >> logn);
  const long long coeff = off & ((1LL << logn) - 1);
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(rec * (1LL << logn) + coeff + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  out[i] = mont_import(z % t, primes[lp.id[limb]]);
}

}  // namespace

hipError_t launch_expand_substitute(const uint64_t* const* cts, int count, int m, uint64_t* S,
                                    int power, int logn, int num_limbs, hipStream_t s) {
  const long long total = ((long long)count * m * 2 * num_limbs) << logn;
  hipLaunchKernelGGL(expand_substitute_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, cts, m,
                     S, power, logn, num_limbs, total);
  return hipGetLastError();
}

hipError_t launch_expand_update(uint64_t* const* cts, int count, int m, const uint64_t* sub0,
                                const uint64_t* sub1, const uint64_t* mono,
                                const PrimeDev* primes, const LimbPrimes& lp, int logn,
                                int num_limbs, hipStream_t s) {
  const long long total = ((long long)count * m * num_limbs) << logn;
  hipLaunchKernelGGL(expand_update_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, cts, m,
                     sub0, sub1, mono, primes, lp, logn, num_limbs, total);
  return hipGetLastError();
}

int pir_k_split(int count, int C, int K, int logn, int num_limbs) {
  const int qg = pir_query_group(count);
  const int cbw = pir_columns_per_thread(qg);
  const long long threads = (long long)((count + qg - 1) / qg) * ((C + cbw - 1) / cbw) *
                            ((long long)num_limbs << logn);
  int ks = 1;
  while (threads * ks < (1LL << 19) && ks * 2 <= 64 && (K + ks * 2 - 1) / (ks * 2) >= 16) ks *= 2;
  return ks;
}

hipError_t launch_pir_inner_product(const uint64_t* const* expd, const uint64_t* expd0,
                                    int count, int K, const uint64_t* db_x, int C,
                                    uint64_t* part, int KS,
                                    uint64_t* const* out, const PrimeDev* primes,
                                    const LimbPrimes& lp, int logn, int num_limbs, hipStream_t s) {
  const long long poly = (long long)num_limbs << logn;
  const int qg = pir_query_group(count);
  const int cbw = pir_columns_per_thread(qg);
  const long long tiles = (poly + kTile - 1) / kTile;
  const long long per_tile = (long long)((count + qg - 1) / qg) * ((C + cbw - 1) / cbw) * KS;
  const dim3 block(kTile);
  if (qg == 4)
    hipLaunchKernelGGL(pir_partial_kernel<4>, dim3((unsigned)(8 * ((tiles + 7) / 8) * per_tile)),
                       block, 0, s, expd, count, K, db_x, C, KS, part, primes, lp, logn,
                       num_limbs);
  else if (qg == 2)
    hipLaunchKernelGGL(pir_partial_kernel<2>, dim3((unsigned)(8 * ((tiles + 7) / 8) * per_tile)),
                       block, 0, s, expd, count, K, db_x, C, KS, part, primes, lp, logn,
                       num_limbs);
  else
    hipLaunchKernelGGL(pir_partial_single_kernel, dim3(grid_of(per_tile * poly, kTile)), block, 0,
                       s, expd0, K, db_x, C, KS, part, primes, lp, logn, num_limbs,
                       per_tile * poly);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long total = (long long)count * C * 2 * poly;
  hipLaunchKernelGGL(pir_sum_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, part, C, KS, out,
                     primes, lp, logn, num_limbs, total);
  return hipGetLastError();
}

hipError_t launch_export_words(const uint64_t* in, uint64_t* out, long long polys,
                               const PrimeDev* primes, const LimbPrimes& lp, int logn,
                               int num_limbs, hipStream_t s) {
  const long long total = (polys * num_limbs) << logn;
  hipLaunchKernelGGL(export_words_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, in, out,
                     primes, lp, logn, num_limbs, total);
  return hipGetLastError();
}

hipError_t launch_db_synth(uint64_t* out, long long records, uint64_t seed, uint64_t t,
                           const PrimeDev* primes, const LimbPrimes& lp, int logn, int num_limbs,
                           hipStream_t s) {
  const long long total = (records * num_limbs) << logn;
  hipLaunchKernelGGL(db_synth_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, out, seed, t,
                     primes, lp, logn, num_limbs, total);
  return hipGetLastError();
}

}  // namespace shg
