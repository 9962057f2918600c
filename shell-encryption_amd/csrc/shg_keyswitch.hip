// Fused gadget key switching for gfx950.
//
// RnsRelinKey::ApplyToRlweCiphertext (rns/rns_relinearization_key.cc:184-232)
// and RnsGaloisKey::ApplyToRlweCiphertext (rns/rns_galois_key.cc:238-287)
// compute, for an input component c in coefficient form,
//   acc0 = init0 + sum_j NTT(d_j) (.) key_b[j],  acc1 = init1 + sum_j NTT(d_j) (.) key_a[j]
// where d_j = g^-1(c) are the unsigned base-2^b digits of the exported
// residues (gadget.h:52-79), lifted to every limb (rns/rns_gadget.cc:125-151).
//
// One workgroup owns one (ciphertext, output limb l).  For each digit j it
// extracts the digit from the source limb straight into VGPRs (no digit
// polynomial is ever written to HBM), imports it modulo q_l, runs the whole
// forward NTT in LDS/VGPRs, and multiplies the transform into two per-thread
// accumulators that stay in VGPRs across all k digits.  Workgroups are
// ordered limb-major so the ciphertexts resident at one time read the same
// key rows, which the L2 and Infinity Cache then serve.
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"
#include "shg_ntt_core.h"

namespace shg {

template <int LOGM, int LOGE, class A>
__global__ void __launch_bounds__((1 << LOGM) >> LOGE)
    keyswitch_kernel(KeySwitchArgs ka, GadgetDesc gd, const PrimeDev* __restrict__ primes,
                     LimbPrimes lp, int count, int num_limbs) {
  using V = typename A::V;
  constexpr int E = 1 << LOGE;
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int N = 1 << LOGM;
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds = reinterpret_cast<V*>(smem);
  const int l = blockIdx.x / count;
  const int ct = blockIdx.x - l * count;
  const int t = threadIdx.x;
  const PrimeDev& P = primes[lp.id[l]];
  const A ar(P);
  const size_t limb_off = (size_t)l << LOGM;
  const size_t poly_words = (size_t)num_limbs << LOGM;

  uint64_t acc0[E], acc1[E];
  uint64_t* o0 = ka.out0 + ct * ka.out_stride + limb_off;
  uint64_t* o1 = ka.out1 + ct * ka.out_stride + limb_off;
  if (ka.accumulate) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      acc0[e] = o0[e * T + t];
      acc1[e] = o1[e * T + t];
    }
  } else {
    const uint64_t* c0 = ka.c0 ? ka.c0 + ct * ka.c0_stride + limb_off : nullptr;
    const uint64_t* c1 = ka.c1 ? ka.c1 + ct * ka.c1_stride + limb_off : nullptr;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      acc0[e] = c0 ? c0[e * T + t] : 0;
      acc1[e] = c1 ? c1[e * T + t] : 0;
    }
  }

  const uint64_t* src_ct = ka.src + ct * ka.src_stride;
  for (int j = 0; j < gd.k; ++j) {
    const int si = gd.src_limb[j];
    const PrimeDev& Ps = primes[lp.id[si]];
    const uint64_t* src = src_ct + ((size_t)si << LOGM);
    const int sh = gd.shift[j];
    const uint64_t mask = (1ull << gd.log_b[j]) - 1;
    V x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint64_t v = mont_export(src[e * T + t], Ps);
      x[e] = ar.load(mont_import((v >> sh) & mask, P));
    }
    FwdPasses<LOGM, LOGE, 0, A>::run(x, lds, t, 0, 0, ar);
    const uint64_t* kb = ka.key_b + (size_t)(ka.key_base + j) * poly_words + limb_off;
    const uint64_t* kav = ka.key_a + (size_t)(ka.key_base + j) * poly_words + limb_off;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = e * T + t;
      const uint64_t d = ar.fwd_out(lds[lpad(p)]);
      acc0[e] = mont_fma(acc0[e], d, kb[p], P);
      acc1[e] = mont_fma(acc1[e], d, kav[p], P);
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    o0[e * T + t] = acc0[e];
    o1[e * T + t] = acc1[e];
  }
  (void)N;
}

template <int LOGM, int LOGE, class A>
static hipError_t launch_ks(const KeySwitchArgs& ka, const GadgetDesc& gd, const PrimeDev* primes,
                            const LimbPrimes& lp, int count, int num_limbs, hipStream_t s) {
  constexpr int T = (1 << LOGM) >> LOGE;
  const size_t shmem = (size_t)lds_words<LOGM>() * 8;
  auto k = keyswitch_kernel<LOGM, LOGE, A>;
  if (shmem > 65536)
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  hipLaunchKernelGGL(k, dim3((unsigned)(count * num_limbs)), dim3(T), shmem, s, ka, gd, primes, lp,
                     count, num_limbs);
  return hipGetLastError();
}

template <class A>
static hipError_t dispatch_ks(int logn, const KeySwitchArgs& ka, const GadgetDesc& gd,
                              const PrimeDev* primes, const LimbPrimes& lp, int count,
                              int num_limbs, hipStream_t s) {
  switch (logn) {
    case 4: return launch_ks<4, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 5: return launch_ks<5, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 6: return launch_ks<6, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 7: return launch_ks<7, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 8: return launch_ks<8, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 9: return launch_ks<9, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 10: return launch_ks<10, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 11: return launch_ks<11, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 12: return launch_ks<12, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 13: return launch_ks<13, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    case 14: return launch_ks<14, 4, A>(ka, gd, primes, lp, count, num_limbs, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_keyswitch(const KeySwitchArgs& ka, const GadgetDesc& gd, int variant,
                            const PrimeDev* primes, const LimbPrimes& lp, int count, int num_limbs,
                            int logn, hipStream_t s) {
  if (variant == SHG_NTT_F64)
    return dispatch_ks<F64Arith>(logn, ka, gd, primes, lp, count, num_limbs, s);
  return dispatch_ks<IntArith>(logn, ka, gd, primes, lp, count, num_limbs, s);
}

}  // namespace shg
