// RNS basis conversions of the BFV side (rns/rns_polynomial.cc:419-701).
//
// For a coefficient-form polynomial a mod Q = q_0...q_{L-1} and another
// chain P = p_0...p_{K-1}, with y_i = [a_i (Q/q_i)^-1]_{q_i}:
//   SwitchRnsBasis          out_j = sum_i y_i (Q/q_i) - u Q            mod p_j
//   ScaleAndSwitchRnsBasis  out_j = u - sum_i y'_i q_i^-1              mod p_j
//                           (y'_i from P * a_i)
//   ApproxSwitchRnsBasis    out_j = sum_i lift(y_i) (Q/q_i)            mod p_j
// where u = round(sum_i y_i / q_i) is formed in double precision exactly as
// the reference forms it (one division and one addition per limb, in limb
// order), so the rounding, and with it every output word, is the
// reference's.  One thread per coefficient reads its L residues once
// (coalesced across the wavefront), keeps y in registers, and writes its K
// outputs: the conversion reads L*8 and writes K*8 bytes per coefficient.
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"

namespace shg {

namespace {

__global__ void basis_switch_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                    BasisArgs ba, const PrimeDev* __restrict__ primes, int logn,
                                    long long total) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const long long n = 1LL << logn;
  const long long k = g & (n - 1);
  const long long poly = g >> logn;
  uint64_t y[kBasisMax];   // Montgomery reps of y_i
  uint64_t yv[kBasisMax];  // their values
  double us = 0.0;
  for (int i = 0; i < ba.nin; ++i) {
    const PrimeDev& Pi = primes[ba.in_id[i]];
    uint64_t a = in[(poly * ba.nin + i) * n + k];
    if (ba.mode == kBasisScale) a = mont_mul(a, ba.pre[i], Pi);
    y[i] = mont_mul(a, ba.q_hat_inv[i], Pi);
    yv[i] = mont_export(y[i], Pi);
    if (ba.mode != kBasisApprox) {
      double s = (double)yv[i];
      s = __ddiv_rn(s, (double)Pi.q);
      us = __dadd_rn(us, s);
    }
  }
  const uint64_t u = ba.mode != kBasisApprox ? (uint64_t)round(us) : 0;
  for (int j = 0; j < ba.nout; ++j) {
    const PrimeDev& Pj = primes[ba.out_id[j]];
    uint64_t acc = 0;
    for (int i = 0; i < ba.nin; ++i) {
      const PrimeDev& Pi = primes[ba.in_id[i]];
      uint64_t c;
      if (ba.balanced)
        c = convert_balanced(y[i], Pi, Pj);                // ConvertModulusBalanced
      else
        c = Pi.q == Pj.q ? y[i] : mont_import(yv[i], Pj);  // ConvertModulus
      acc = mod_add(acc, mont_mul(c, ba.mat[i * ba.nout + j], Pj), Pj.q);
    }
    if (ba.mode == kBasisExact)
      acc = mod_sub(acc, mont_mul(mont_import(u, Pj), ba.vec[j], Pj), Pj.q);
    else if (ba.mode == kBasisScale)
      acc = mod_add(acc, mont_import(u, Pj), Pj.q);
    out[(poly * ba.nout + j) * n + k] = acc;
  }
}

}  // namespace

hipError_t launch_basis_switch(const uint64_t* in, uint64_t* out, const BasisArgs& ba,
                               const PrimeDev* primes, int count, int logn, hipStream_t s) {
  const long long total = (long long)count << logn;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(basis_switch_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     in, out, ba, primes, logn, total);
  return hipGetLastError();
}

}  // namespace shg
