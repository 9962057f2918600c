// Persistent launch over uneven limb jobs.
//
// A job is one limb of n = 2^log_n words (2^11 <= n <= 2^15) at its own
// address with its own prime: the forward or inverse negacyclic transform
// (ForwardNumberTheoreticTransform / InverseNumberTheoreticTransform,
// shell_encryption/dft_transformations.h:158-201) or the Montgomery product
// with a second limb (MontgomeryInt::MulInPlace, montgomery.h:318-336).  One
// launch of at most one workgroup per resident slot runs every job: the
// workgroup (1024 threads) loops over jobs blockIdx.x, blockIdx.x +
// gridDim.x, ... of a table sorted by decreasing size, so the static stride
// gives each workgroup a similar mix.  The launch needs no host work beyond
// the kernel itself, so a chain of such launches can be captured in a HIP
// graph.
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"
#include "shg_ntt_core.h"

namespace shg {

namespace {

constexpr int kT = kBatchThreads;

// Forward transform of the 2^LOGM-word slice h of a limb with LOGM + HB
// bits (the first HB stages already applied), in place at g.
template <int LOGM, int HB, class A>
__device__ __forceinline__ void fwd_slice(gmut<uint64_t> g, int h, const A& ar, typename A::V* lds,
                                          int t) {
  constexpr int LOGE = LOGM - 10;
  constexpr int E = 1 << LOGE;
  asm volatile("" : "+v"(t));  // keep this body's addressing local to it
  typename A::V x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = ar.load(g[e * kT + t]);
  FwdPasses<LOGM, LOGE, 0, A, true, HB>::run(x, lds, t, h, ar);
#pragma unroll
  for (int e = 0; e < E; ++e) g[e * kT + t] = ar.fwd_out(lds[lpad(e * kT + t)]);
  // The next job's first exchange overwrites what was just read.
  __syncthreads();
}

// Inverse transform of slice h (local stages only), outputs times n^-1 when
// scale is 1, canonical when 0.
template <int LOGM, int HB, class A>
__device__ __forceinline__ void inv_slice(gmut<uint64_t> g, int h, const A& ar, typename A::V* lds,
                                          int scale, int t) {
  using V = typename A::V;
  constexpr int LOGE = LOGM - 10;
  constexpr int E = 1 << LOGE;
  asm volatile("" : "+v"(t));  // keep this body's addressing local to it
#pragma unroll
  for (int e = 0; e < E; ++e) lds[lpad(e * kT + t)] = ar.load(g[e * kT + t]);
  __syncthreads();
  V x[E];
  InvPasses<LOGM, LOGE, 0, A, false, LOGM + HB>::run(x, lds, t, h, ar);
  constexpr int I0 = InvLast<LOGM, LOGE, 0, A>::value;
  constexpr int R = LOGM - I0;
  constexpr int G = E >> R;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
#pragma unroll
    for (int e = 0; e < (1 << R); ++e)
      g[(e << I0) | (gg * kT + t)] = ar.inv_out(x[gg * (1 << R) + e], scale);
  }
  // The last pass read its exchange before its own barrier: the next job
  // may write LDS at once.
}

template <class A>
__device__ __forceinline__ void fwd_job(gmut<uint64_t> g, int log_n, const A& ar, typename A::V* lds,
                                        int t) {
  switch (log_n) {
    case 11: fwd_slice<11, 0>(g, 0, ar, lds, t); break;
    case 12: fwd_slice<12, 0>(g, 0, ar, lds, t); break;
    case 13: fwd_slice<13, 0>(g, 0, ar, lds, t); break;
    case 14: fwd_slice<14, 0>(g, 0, ar, lds, t); break;
    case 15: {
      // Stage 0 over the whole limb (pairs i, i + 2^14), then the two
      // halves as 2^14-word slices.  __syncthreads makes the global writes
      // of every thread visible to the whole workgroup.
      constexpr int H = 1 << 14;
      const auto w = ar.fwd_tw(1);
      for (int i = t; i < H; i += kT) {
        auto u = ar.load(g[i]);
        auto v = ar.load(g[i + H]);
        ar.ct(u, v, w);
        g[i] = ar.fwd_out(u);
        g[i + H] = ar.fwd_out(v);
      }
      __syncthreads();
      fwd_slice<14, 1>(g, 0, ar, lds, t);
      fwd_slice<14, 1>(g + H, 1, ar, lds, t);
      break;
    }
    default: break;
  }
}

template <class A>
__device__ __forceinline__ void inv_job(gmut<uint64_t> g, int log_n, const A& ar, typename A::V* lds,
                                        int t) {
  switch (log_n) {
    case 11: inv_slice<11, 0>(g, 0, ar, lds, 1, t); break;
    case 12: inv_slice<12, 0>(g, 0, ar, lds, 1, t); break;
    case 13: inv_slice<13, 0>(g, 0, ar, lds, 1, t); break;
    case 14: inv_slice<14, 0>(g, 0, ar, lds, 1, t); break;
    case 15: {
      // Stages 0..13 on each half, then the last Gentleman-Sande stage
      // (pairs i, i + 2^14, twiddle index n - 2) with the n^-1 scaling.
      constexpr int H = 1 << 14;
      inv_slice<14, 1>(g, 0, ar, lds, 0, t);
      inv_slice<14, 1>(g + H, 1, ar, lds, 0, t);
      __syncthreads();
      const auto w = ar.inv_tw(2 * H - 2);
      for (int i = t; i < H; i += kT) {
        auto u = ar.load(g[i]);
        auto v = ar.load(g[i + H]);
        ar.gs(u, v, w);
        g[i] = ar.inv_out(u, 1);
        g[i + H] = ar.inv_out(v, 1);
      }
      // The next job's LDS writes are independent of these global ones.
      break;
    }
    default: break;
  }
}

template <class A>
__global__ void __launch_bounds__(kT)
    batch_limbs_kernel(const LimbJobDev* __restrict__ jobs, int num_jobs, int op,
                       const PrimeDev* __restrict__ primes) {
  extern __shared__ __align__(16) unsigned char smem[];
  auto* lds = reinterpret_cast<typename A::V*>(smem);
  for (int j = blockIdx.x; j < num_jobs; j += gridDim.x) {
    // An opaque copy of the lane index per job keeps the address arithmetic
    // of the ten transform bodies inside the loop: hoisted out of it, their
    // loop-invariant offsets would all stay live at once and spill.
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const LimbJobDev job = jobs[j];
    const PrimeDev& P = primes[job.prime];
    if (op == kBatchMul) {
      const int n = 1 << job.log_n;
      gptr<u64x2> b = (gptr<u64x2>)job.other;
      gmut<u64x2> a = (gmut<u64x2>)job.data;
      for (int i = t; i < n / 2; i += kT) {
        u64x2 x = a[i];
        const u64x2 y = b[i];
        x.x = mont_mul(x.x, y.x, P);
        x.y = mont_mul(x.y, y.y, P);
        a[i] = x;
      }
      continue;
    }
    const A ar(P);
    const gmut<uint64_t> g = (gmut<uint64_t>)job.data;
    if (op == kBatchNttFwd)
      fwd_job(g, job.log_n, ar, lds, t);
    else
      inv_job(g, job.log_n, ar, lds, t);
  }
}

}  // namespace

size_t batch_shmem_bytes(int max_log_n) {
  const int logm = max_log_n < 14 ? max_log_n : 14;
  return (size_t)((1 << logm) + ((1 << logm) >> 4)) * 8;
}

const void* batch_kernel_ptr(int variant) {
  return variant == SHG_NTT_F64 ? (const void*)batch_limbs_kernel<F64Arith>
                                : (const void*)batch_limbs_kernel<IntArith>;
}

hipError_t launch_batch_limbs(const LimbJobDev* jobs, int num_jobs, int op, int variant,
                              const PrimeDev* primes, int grid, size_t shmem, hipStream_t s) {
  if (num_jobs <= 0) return hipSuccess;
  if (variant == SHG_NTT_F64)
    hipLaunchKernelGGL(batch_limbs_kernel<F64Arith>, dim3(grid), dim3(kT), shmem, s, jobs,
                       num_jobs, op, primes);
  else
    hipLaunchKernelGGL(batch_limbs_kernel<IntArith>, dim3(grid), dim3(kT), shmem, s, jobs,
                       num_jobs, op, primes);
  return hipGetLastError();
}

}  // namespace shg
