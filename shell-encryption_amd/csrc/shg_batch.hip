// Persistent launch over uneven limb jobs.
//
// A job is one limb of n = 2^log_n words (2^11 <= n <= 2^15) at its own
// address with its own prime: the forward or inverse negacyclic transform
// (ForwardNumberTheoreticTransform / InverseNumberTheoreticTransform,
// shell_encryption/dft_transformations.h:158-201) or the Montgomery product
// with a second limb (MontgomeryInt::MulInPlace, montgomery.h:318-336); or,
// per polynomial, ModReduceLsb.  One
// launch of at most one workgroup per resident slot runs every job: the
// workgroup (1024 threads) loops over jobs blockIdx.x, blockIdx.x +
// gridDim.x, ... of a table sorted by decreasing size, so the static stride
// gives each workgroup a similar mix.  The launch needs no host work beyond
// the kernel itself, so a chain of such launches can be captured in a HIP
// graph.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "shg_device.h"
#include "shg_kernels.h"
#include "shg_ntt_core.h"
#include "shg_ntt_wave.h"
#include "shg_ntt_w32.h"

// 2^14-word slices of FP64 jobs on the wave-exchange layout (shg_ntt_wave.h).
#ifndef SHG_BATCH_WAVE
#define SHG_BATCH_WAVE 1
#endif

namespace shg {

namespace {

constexpr int kT = kBatchThreads;

// Forward transform of the 2^LOGM-word slice h of a limb with LOGM + HB
// bits (the first HB stages already applied).  ld(p) gives the input word
// at slice position p (already in the policy's representation), st(p, v)
// takes the canonical output word.
template <int LOGM, int HB, class A, class Ld, class St>
__device__ __forceinline__ void fwd_slice_io(int h, const A& ar, typename A::V* lds, int t,
                                             const Ld& ld, const St& st) {
  if constexpr (LOGM == 14 && std::is_same<A, F64Arith>::value && SHG_BATCH_WAVE) {
    wave_fwd_slice<HB>(h, ar, lds, ld, st);
    return;
  }
  constexpr int LOGE = LOGM - 10;
  constexpr int E = 1 << LOGE;
  asm volatile("" : "+v"(t));  // keep this body's addressing local to it
  typename A::V x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = ld(e * kT + t);
  FwdPasses<LOGM, LOGE, 0, A, true, HB>::run(x, lds, t, h, ar);
  // The next job's first exchange overwrites what is read here: each group
  // of (up to) four reads is pinned before the barrier.
  constexpr int C = E < 4 ? E : 4;
#pragma unroll
  for (int e0 = 0; e0 < E; e0 += C) {
    typename A::V y[C];
#pragma unroll
    for (int e = 0; e < C; ++e) y[e] = lds[lpad((e0 + e) * kT + t)];
    pin_words(y);
#pragma unroll
    for (int e = 0; e < C; ++e) st((e0 + e) * kT + t, ar.fwd_out(y[e]));
  }
  __syncthreads();
}

// Inverse transform of slice h (local stages only), outputs times n^-1 when
// scale is 1, canonical when 0; ld / st as for fwd_slice_io (ld returns the
// raw residue word).
template <int LOGM, int HB, class A, class Ld, class St>
__device__ __forceinline__ void inv_slice_io(int h, const A& ar, typename A::V* lds, int scale,
                                             int t, const Ld& ld, const St& st) {
  if constexpr (LOGM == 14 && std::is_same<A, F64Arith>::value && SHG_BATCH_WAVE) {
    wave_inv_slice<HB>(h, ar, lds, scale, ld, st);
    return;
  }
  using V = typename A::V;
  constexpr int LOGE = LOGM - 10;
  constexpr int E = 1 << LOGE;
  asm volatile("" : "+v"(t));  // keep this body's addressing local to it
#pragma unroll
  for (int e = 0; e < E; ++e) lds[lpad(e * kT + t)] = ar.load(ld(e * kT + t));
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
      st((e << I0) | (gg * kT + t), ar.inv_out(x[gg * (1 << R) + e], scale));
  }
  // The last pass read its exchange before its own barrier: the next job
  // may write LDS at once.
}

// Forward transform of one limb of 2^log_n words whose input word p is ld(p)
// (policy representation) and whose output word p goes to st(p, v).  For
// n = 2^15 stage 0 runs over the whole limb into tmp (2^15 words; may be
// the limb itself when ld reads it), then the halves run as 2^14-word
// slices; __syncthreads makes every thread's global writes visible to the
// whole workgroup.
//
// HOLD (FP64 wave path): for n = 2^15 each thread reads the pairs (p, p + H)
// of its pass-1 positions, applies stage 0, runs half 0 from the sums and
// keeps the differences in VGPRs for half 1: the limb is read and written
// once, with no pass through tmp.
template <bool HOLD = false, class A, class Ld, class St>
__device__ __forceinline__ void fwd_io(int log_n, const A& ar, typename A::V* lds, int t,
                                       gmut<uint64_t> tmp, const Ld& ld, const St& st) {
  switch (log_n) {
    case 11: fwd_slice_io<11, 0>(0, ar, lds, t, ld, st); break;
    case 12: fwd_slice_io<12, 0>(0, ar, lds, t, ld, st); break;
    case 13: fwd_slice_io<13, 0>(0, ar, lds, t, ld, st); break;
    case 14: fwd_slice_io<14, 0>(0, ar, lds, t, ld, st); break;
    case 15: {
      constexpr int H = 1 << 14;
      if constexpr (HOLD && std::is_same<A, F64Arith>::value && SHG_BATCH_WAVE) {
        using Lay = WaveLayout<4>;
        const auto w = ar.fwd_tw(1);
        const int tp1 = Lay::thread_part(1, (int)threadIdx.x & 63,
                                         __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
        double x[16], held[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int p = tp1 + Lay::reg_part(1, e);
          double u = ld(p);
          double v = ld(p + H);
          ar.ct(u, v, w);
          x[e] = u;
          held[e] = v;
        }
        auto none = [](int) { return 0.0; };
        // Every thread reads its inputs before the core's first barrier, so
        // half 0's stores cannot reach a word that is still to be read.
        const int tp4 = wave_fwd_core<1, true>(0, ar, lds, none, x);
#pragma unroll
        for (int e = 0; e < 16; ++e) st(tp4 + Lay::reg_part(4, e), ar.fwd_out(x[e]));
        __syncthreads();
        wave_fwd_core<1, true>(1, ar, lds, none, held);
#pragma unroll
        for (int e = 0; e < 16; ++e) st(H + tp4 + Lay::reg_part(4, e), ar.fwd_out(held[e]));
        __syncthreads();
        break;
      }
      const auto w = ar.fwd_tw(1);
      for (int i = t; i < H; i += kT) {
        auto u = ld(i);
        auto v = ld(i + H);
        ar.ct(u, v, w);
        tmp[i] = ar.fwd_out(u);
        tmp[i + H] = ar.fwd_out(v);
      }
      __syncthreads();
      fwd_slice_io<14, 1>(0, ar, lds, t, [&](int p) { return ar.load(tmp[p]); }, st);
      fwd_slice_io<14, 1>(1, ar, lds, t, [&](int p) { return ar.load(tmp[H + p]); },
                          [&](int p, uint64_t v) { st(H + p, v); });
      break;
    }
    default: break;
  }
}

// Inverse transform (times n^-1) of one limb whose input word p is ld(p)
// (raw residue) and whose output goes to st(p, v).  For n = 2^15 the halves
// run as slices into tmp, then the last Gentleman-Sande stage (pairs i,
// i + 2^14, twiddle index n - 2) applies the scaling.
//
// HOLD (FP64 wave path): for n = 2^15 both halves' slices end with their
// words in VGPRs at the same positions, so the last stage (pairs i,
// i + 2^14) runs in registers with the n^-1 scaling folded in, and the limb
// is read and written once.
template <bool HOLD = false, class A, class Ld, class St>
__device__ __forceinline__ void inv_io(int log_n, const A& ar, typename A::V* lds, int t,
                                       gmut<uint64_t> tmp, const Ld& ld, const St& st) {
  switch (log_n) {
    case 11: inv_slice_io<11, 0>(0, ar, lds, 1, t, ld, st); break;
    case 12: inv_slice_io<12, 0>(0, ar, lds, 1, t, ld, st); break;
    case 13: inv_slice_io<13, 0>(0, ar, lds, 1, t, ld, st); break;
    case 14: inv_slice_io<14, 0>(0, ar, lds, 1, t, ld, st); break;
    case 15: {
      constexpr int H = 1 << 14;
      if constexpr (HOLD && std::is_same<A, F64Arith>::value && SHG_BATCH_WAVE) {
        using Lay = WaveLayout<4>;
        double y0[16], y1[16], tw4[15];
        const int tp1 = wave_inv_core<1>(0, ar, lds, ld, y0, tw4);
        wave_pass_inv<4, 1, 3, 0>(y0, tw4, ar);
        __syncthreads();
        wave_inv_core<1>(1, ar, lds, [&](int p) { return ld(H + p); }, y1, tw4);
        wave_pass_inv<4, 1, 3, 0>(y1, tw4, ar);
        __syncthreads();
        // Stage on index bit 14 (twiddle index 2n' - 2 of the 2^15 table)
        // with the n^-1 scale: (u + v) s, (u - v)(w s).  Inputs are four
        // Gentleman-Sande stages past the reduction at bit 10, below 8q.
        const double s = ar.ninv;
        const double ws = f64_mulmod_h(ar.inv_tw_raw(2 * H - 2), s, ar.q, ar.qinv);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int p = tp1 + Lay::reg_part(1, e);
          st(p, ar.fwd_out(f64_mulmod_h(y0[e] + y1[e], s, ar.q, ar.qinv)));
          st(H + p, ar.fwd_out(f64_mulmod_h(y0[e] - y1[e], ws, ar.q, ar.qinv)));
        }
        break;
      }
      inv_slice_io<14, 1>(0, ar, lds, 0, t, ld, [&](int p, uint64_t v) { tmp[p] = v; });
      inv_slice_io<14, 1>(1, ar, lds, 0, t, [&](int p) { return ld(H + p); },
                          [&](int p, uint64_t v) { tmp[H + p] = v; });
      __syncthreads();
      const auto w = ar.inv_tw(2 * H - 2);
      for (int i = t; i < H; i += kT) {
        auto u = ar.load(tmp[i]);
        auto v = ar.load(tmp[i + H]);
        ar.gs(u, v, w);
        st(i, ar.inv_out(u, 1));
        st(i + H, ar.inv_out(v, 1));
      }
      break;
    }
    default: break;
  }
}

template <class A>
__device__ __forceinline__ void fwd_job(gmut<uint64_t> g, int log_n, const A& ar,
                                        typename A::V* lds, int t) {
  fwd_io<true>(log_n, ar, lds, t, g, [&](int p) { return ar.load(g[p]); },
               [&](int p, uint64_t v) { g[p] = v; });
}

template <class A>
__device__ __forceinline__ void inv_job(gmut<uint64_t> g, int log_n, const A& ar,
                                        typename A::V* lds, int t) {
  inv_io<true>(log_n, ar, lds, t, g, [&](int p) { return (uint64_t)g[p]; },
               [&](int p, uint64_t v) { g[p] = v; });
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

// ModReduceLsb (rns/rns_polynomial.cc:262-344) of one polynomial per job,
// NTT or coefficient form: delta = INTT(k * a_l) (coefficient form, k =
// (q_l - (q_l mod t)) / t) goes to this workgroup's scratch, then every
// other limb i becomes (r * a_i + t * NTT(lift_i(delta))) * q_l^-1 with the
// balanced lift of ConvertModulusBalanced (modulus_conversion.h:30-83) and
// r = q_l mod t.  Limbs 0..L-2 hold the result; the last limb is left as
// it was.  Each workgroup owns 2^16 words of scratch.  UNIFORM: every job
// is `uni` with its data at uni.data + j * stride (one call over equal
// polynomials; the job then lives in the kernel arguments).
template <class A, bool UNIFORM>
__global__ void __launch_bounds__(kT)
    batch_modreduce_kernel(const PolyJobDev* __restrict__ jobs, int num_jobs, PolyJobDev uni,
                           size_t stride, const PrimeDev* __restrict__ primes,
                           uint64_t* __restrict__ scratch) {
  extern __shared__ __align__(16) unsigned char smem[];
  auto* lds = reinterpret_cast<typename A::V*>(smem);
  const gmut<uint64_t> s0 = (gmut<uint64_t>)(scratch + ((size_t)blockIdx.x << 16));
  const gmut<uint64_t> s1 = s0 + (1 << 15);
  for (int j = blockIdx.x; j < num_jobs; j += gridDim.x) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const PolyJobDev& job = UNIFORM ? uni : jobs[j];
    const int L = job.num_limbs;
    const int log_n = job.log_n;
    const int n = 1 << log_n;
    const gmut<uint64_t> a = (gmut<uint64_t>)(UNIFORM ? uni.data + (size_t)j * stride : job.data);
    const PrimeDev& Pl = primes[job.prime[L - 1]];
    const gmut<uint64_t> al = a + (size_t)(L - 1) * n;
    const uint64_t k = job.k_mod_ql;
    if constexpr (std::is_same<A, F64Arith>::value) {
      if (job.is_ntt) {
        // FP64 form: delta = exported INTT(k * a_l), balanced, kept as a
        // double; each other limb is a_i * c1 + NTT(delta mod q_i) * c2.
        A arl(Pl);
        arl.ninv = job.f_kl;
        arl.ninvq = job.f_kl * arl.qinv;
        const uint64_t half_ql = Pl.q >> 1;
        inv_io<true>(log_n, arl, lds, t, s0, [&](int p) { return (uint64_t)al[p]; },
               [&](int p, uint64_t v) {
                 const double d = u52_to_f64(v) - (v > half_ql ? arl.q : 0.0);
                 s0[p] = (uint64_t)__double_as_longlong(d);
               });
        __syncthreads();
        for (int i = 0; i < L - 1; ++i) {
          const A ari(primes[job.prime[i]]);
          const gmut<uint64_t> ai = a + (size_t)i * n;
          const double c1 = job.f_c1[i], c1q = c1 * ari.qinv;
          const double c2 = job.f_c2[i], c2q = c2 * ari.qinv;
          fwd_io<true>(log_n, ari, lds, t, s1,
                 [&](int p) {
                   return f64_reduce(__longlong_as_double((long long)(uint64_t)s0[p]), ari.q,
                                     ari.qinv);
                 },
                 [&](int p, uint64_t v) {
                   const double sum = f64_mulmod(u52_to_f64(ai[p]), c1, c1q, ari.q) +
                                      f64_mulmod(u52_to_f64(v), c2, c2q, ari.q);
                   ai[p] = f64_canon_u52(f64_reduce(sum, ari.q, ari.qinv), ari.q);
                 });
        }
        // delta (s0) is rewritten by the next job.
        __syncthreads();
        continue;
      }
    }
    if (job.is_ntt) {
      const A arl(Pl);
      inv_io(log_n, arl, lds, t, s0, [&](int p) { return mont_mul(al[p], k, Pl); },
             [&](int p, uint64_t v) { s0[p] = v; });
    } else {
      for (int p = t; p < n; p += kT) s0[p] = mont_mul(al[p], k, Pl);
    }
    __syncthreads();
    for (int i = 0; i < L - 1; ++i) {
      const PrimeDev& Pi = primes[job.prime[i]];
      const gmut<uint64_t> ai = a + (size_t)i * n;
      const uint64_t r = job.r_mod_qi[i], tq = job.t_mod_qi[i], qinv = job.ql_inv[i];
      auto combine = [&](int p, uint64_t d) {
        ai[p] = mont_mul(mod_add(mont_mul(ai[p], r, Pi), mont_mul(d, tq, Pi), Pi.q), qinv, Pi);
      };
      if (job.is_ntt) {
        const A ari(Pi);
        fwd_io(log_n, ari, lds, t, s1, [&](int p) { return ari.load(convert_balanced(s0[p], Pl, Pi)); },
               combine);
      } else {
        for (int p = t; p < n; p += kT) combine(p, convert_balanced(s0[p], Pl, Pi));
      }
    }
    // delta (s0) is rewritten by the next job.
    __syncthreads();
  }
}

// Whole 2^15 limbs with FP64 arithmetic, persistent over the limbs: the
// register-held forms of fwd_io / inv_io (a limb crosses HBM once each way,
// the inverse scaled by n^-1).  Limb `item` uses prime lp.id[item % num_limbs].
__global__ void __launch_bounds__(kT)
    ntt15_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                 LimbPrimes lp, int num_limbs, int items, int inverse) {
  extern __shared__ __align__(16) unsigned char smem[];
  auto* lds = reinterpret_cast<double*>(smem);
  for (int item = blockIdx.x; item < items; item += gridDim.x) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const F64Arith ar(primes[lp.id[item % num_limbs]]);
    const gmut<uint64_t> g = (gmut<uint64_t>)(data + ((size_t)item << 15));
    if (inverse)
      inv_job(g, 15, ar, lds, t);
    else
      fwd_job(g, 15, ar, lds, t);
  }
}

}  // namespace

hipError_t launch_ntt15(bool inverse, uint64_t* data, const PrimeDev* primes,
                        const LimbPrimes& lp, int rows, int num_limbs, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  const size_t shmem = batch_shmem_bytes(15);
  const void* k = (const void*)ntt15_kernel;
  set_max_dynamic_shmem(k, shmem);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kT, shmem) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  long long grid = (long long)cus * per_cu;
  // A multiple of the limb count keeps each workgroup on one prime.
  if (grid >= num_limbs) grid -= grid % num_limbs;
  if (grid > rows) grid = rows;
  hipLaunchKernelGGL(ntt15_kernel, dim3((unsigned)grid), dim3(kT), shmem, s, data, primes, lp,
                     num_limbs, rows, inverse ? 1 : 0);
  return hipGetLastError();
}

size_t batch_shmem_bytes(int max_log_n) {
  const int logm = max_log_n < 14 ? max_log_n : 14;
  return (size_t)((1 << logm) + ((1 << logm) >> 4)) * 8;
}

const void* batch_kernel_ptr(int variant, bool modreduce) {
  if (modreduce)
    return variant == SHG_NTT_F64 ? (const void*)batch_modreduce_kernel<F64Arith, false>
                                  : (const void*)batch_modreduce_kernel<IntArith, false>;
  return variant == SHG_NTT_F64 ? (const void*)batch_limbs_kernel<F64Arith>
                                : (const void*)batch_limbs_kernel<IntArith>;
}

hipError_t launch_batch_modreduce(const PolyJobDev* jobs, int num_jobs, int variant,
                                  const PrimeDev* primes, uint64_t* scratch, int grid,
                                  size_t shmem, hipStream_t s) {
  if (num_jobs <= 0) return hipSuccess;
  const PolyJobDev none{};
  if (variant == SHG_NTT_F64)
    hipLaunchKernelGGL((batch_modreduce_kernel<F64Arith, false>), dim3(grid), dim3(kT), shmem, s,
                       jobs, num_jobs, none, (size_t)0, primes, scratch);
  else
    hipLaunchKernelGGL((batch_modreduce_kernel<IntArith, false>), dim3(grid), dim3(kT), shmem, s,
                       jobs, num_jobs, none, (size_t)0, primes, scratch);
  return hipGetLastError();
}

static const void* modreduce_uniform_kernel(int variant) {
  return variant == SHG_NTT_F64 ? (const void*)batch_modreduce_kernel<F64Arith, true>
                                : (const void*)batch_modreduce_kernel<IntArith, true>;
}

int modreduce_uniform_grid(int variant, int log_n, int count) {
  const void* k = modreduce_uniform_kernel(variant);
  const size_t shmem = batch_shmem_bytes(log_n);
  if (shmem > 65536) set_max_dynamic_shmem(k, shmem);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kT, shmem) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  const long long cap = (long long)cus * per_cu;
  return (int)(cap < count ? cap : count);
}

hipError_t launch_modreduce_uniform(const PolyJobDev& uni, size_t stride, int count, int variant,
                                    const PrimeDev* primes, uint64_t* scratch, int grid,
                                    hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const size_t shmem = batch_shmem_bytes(uni.log_n);
  if (variant == SHG_NTT_F64)
    hipLaunchKernelGGL((batch_modreduce_kernel<F64Arith, true>), dim3(grid), dim3(kT), shmem, s,
                       nullptr, count, uni, stride, primes, scratch);
  else
    hipLaunchKernelGGL((batch_modreduce_kernel<IntArith, true>), dim3(grid), dim3(kT), shmem, s,
                       nullptr, count, uni, stride, primes, scratch);
  return hipGetLastError();
}

// ModReduceLsb of `count` equal polynomials in NTT form at n = 2^14 with
// FP64 arithmetic, on the two-workgroups-per-CU transforms (shg_ntt_w32.h)
// in two launches; the same operations as batch_modreduce_kernel's FP64
// path, so the same words.
// Launch 1, one workgroup per polynomial: delta = the inverse transform of
// a_l scaled by k n^-1 2^-64 (the exported k a_l, canonical), lifted to the
// balanced representative of ConvertModulusBalanced (v > q_l / 2 -> v - q_l)
// and stored in scratch as a double.
struct W32ModDelta {
  __amdgpu_buffer_rsrc_t a, s;
  const F64Arith* ar;
  uint64_t half;
  static constexpr bool kAll = false;
  __device__ __forceinline__ void store_all(int, const double (&)[32]) const {}
  __device__ __forceinline__ void load(int vs, int k, double& lo, double& hi) const {
    const U64Pair v = w32::ld128(a, vs, k * 1024);
    lo = ar->load(v.x);
    hi = ar->load(v.y);
  }
  __device__ __forceinline__ void store(int v1, int e, double x) const {
    const uint64_t v = ar->fwd_out(x);
    const double d = u52_to_f64(v) - (v > half ? ar->q : 0.0);
    w32::st64(s, (uint64_t)__double_as_longlong(d), v1, e * 4096);
  }
};
// Launch 2, one workgroup per (polynomial, limb i < L - 1): the forward
// transform mod q_i of delta mod q_i, then a_i <- a_i c1[i] + NTT c2[i]
// (exact products; c1 = r q_l^-1 2^-128, c2 = t q_l^-1 2^-64).
struct W32ModLimb {
  __amdgpu_buffer_rsrc_t s, a;
  const F64Arith* ar;
  double c1, c1q, c2, c2q;
  static constexpr bool kAux = true;  // a_i is read before the store
  using Aux = U64Pair;
  __device__ __forceinline__ Aux aux(int vs, int k) const { return w32::ld128(a, vs, k * 1024); }
  __device__ __forceinline__ double load(int v1, int e) const {
    const double d = __longlong_as_double((long long)w32::ld64(s, v1, e * 4096));
    return f64_reduce(d, ar->q, ar->qinv);
  }
  __device__ __forceinline__ uint64_t combine(uint64_t aw, double z) const {
    const double sum = f64_mulmod(u52_to_f64(aw), c1, c1q, ar->q) + f64_mulmod(z, c2, c2q, ar->q);
    return f64_canon_u52(f64_reduce(sum, ar->q, ar->qinv), ar->q);
  }
  __device__ __forceinline__ void store(int vs, int k, double lo, double hi, Aux ai) const {
    U64Pair out;
    out.x = combine(ai.x, lo);
    out.y = combine(ai.y, hi);
    w32::st128(a, out, vs, k * 1024);
  }
};

__global__ void __launch_bounds__(w32::kThreads, 4)
    modreduce_w32_delta_kernel(PolyJobDev uni, size_t stride, const PrimeDev* __restrict__ primes,
                               uint64_t* __restrict__ scratch) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = blockIdx.x;
  const int L = uni.num_limbs;
  const PrimeDev& Pl = primes[uni.prime[L - 1]];
  const F64Arith ar(Pl);
  const W32ModDelta io{w32::rsrc(uni.data + (size_t)p * stride + ((size_t)(L - 1) << w32::kLogN)),
                       w32::rsrc(scratch + ((size_t)p << w32::kLogN)), &ar, Pl.q >> 1};
  w32::inv_limb(ar, Pl, reinterpret_cast<double*>(smem), uni.f_kl, io);
}

// Workgroups are remapped so that each XCD runs one contiguous range of
// (polynomial, limb) items: the L - 1 limbs of a polynomial then read its
// delta through one L2.
__global__ void __launch_bounds__(w32::kThreads, 4)
    modreduce_w32_limb_kernel(PolyJobDev uni, size_t stride, const PrimeDev* __restrict__ primes,
                              const uint64_t* __restrict__ scratch) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned b = blockIdx.x;
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);
  const int nr = uni.num_limbs - 1;
  const int p = (int)b / nr;
  const int i = (int)b - p * nr;
  const PrimeDev& Pi = primes[uni.prime[i]];
  const F64Arith ar(Pi);
  const double c1 = uni.f_c1[i], c2 = uni.f_c2[i];
  const W32ModLimb io{w32::rsrc(scratch + ((size_t)p << w32::kLogN)),
                      w32::rsrc(uni.data + (size_t)p * stride + ((size_t)i << w32::kLogN)), &ar,
                      c1, c1 * ar.qinv, c2, c2 * ar.qinv};
  w32::fwd_limb(ar, Pi, reinterpret_cast<double*>(smem), io);
}

hipError_t launch_modreduce_w32(const PolyJobDev& uni, size_t stride, int count,
                                const PrimeDev* primes, uint64_t* scratch, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const size_t shmem = (size_t)w32::kLdsWords * 8;
  set_max_dynamic_shmem((const void*)modreduce_w32_delta_kernel, shmem);
  set_max_dynamic_shmem((const void*)modreduce_w32_limb_kernel, shmem);
  hipLaunchKernelGGL(modreduce_w32_delta_kernel, dim3(count), dim3(w32::kThreads), shmem, s, uni,
                     stride, primes, scratch);
  hipLaunchKernelGGL(modreduce_w32_limb_kernel, dim3(count * (uni.num_limbs - 1)),
                     dim3(w32::kThreads), shmem, s, uni, stride, primes,
                     (const uint64_t*)scratch);
  return hipGetLastError();
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
