// Fused gadget key switching for gfx950.
//
// RnsRelinKey::ApplyToRlweCiphertext (rns/rns_relinearization_key.cc:184-232)
// and RnsGaloisKey::ApplyToRlweCiphertext (rns/rns_galois_key.cc:238-287)
// compute, for an input component c in coefficient form,
//   acc0 = init0 + sum_j NTT(d_j) (.) key_b[j],  acc1 = init1 + sum_j NTT(d_j) (.) key_a[j]
// where d_j = g^-1(c) are the unsigned base-2^b digits of the exported
// residues (gadget.h:52-79), lifted to every limb (rns/rns_gadget.cc:125-151).
//
// One workgroup owns one (ciphertext, output limb l) (two for n = 2^14).  The
// source component arrives already exported (the inverse transform that
// produced it folds ExportInt into its n^-1 scaling).  For each digit j the
// kernel extracts the digit from the source limb straight into VGPRs (no
// digit polynomial is ever written to HBM), runs the whole forward NTT in
// LDS/VGPRs, and multiplies the transform into two per-thread accumulators
// that stay in VGPRs across all k digits.  Workgroups are
// ordered limb-major so the ciphertexts resident at one time read the same
// key rows, which the L2 and Infinity Cache then serve.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "shg_device.h"
#include "shg_kernels.h"
#include "shg_ntt_core.h"
#include "shg_ntt_wave.h"
#include "shg_ntt_w32.h"

#ifndef SHG_KS_XCD_REMAP
#define SHG_KS_XCD_REMAP 1
#endif
#ifndef SHG_KS_DB
#define SHG_KS_DB 1
#endif
#ifndef SHG_KS_WAVE
#define SHG_KS_WAVE 1
#endif
// Key rows read through buffer loads (descriptor per digit in SGPRs, one
// 32-bit lane offset shared by every digit): 10.27-10.29 vs 10.44-10.49 ms
// per 512 ciphertexts at n = 2^14, L = 8, together with the stage-0 FMA.
// Key words requested per batch ahead of their products (16-byte pairs of
// each key polynomial): at 2^14 with 8 words per thread, 2 pairs measured
// 8.63-8.64 ms per 512 ciphertexts, 1 pair 8.80, one load per product 8.65;
// 4 pairs spill.
#ifndef SHG_KS_KEY_BATCH
#define SHG_KS_KEY_BATCH 2
#endif
#ifndef SHG_KS_KEY_BUF
#define SHG_KS_KEY_BUF 1
#endif
// Stage 0 of a split limb from a table of its products: with digits of at
// most kKsTabLogB bits, the product w * v that the first Cooley-Tukey stage
// forms for each digit v of the other half takes one of 2^b values, held in
// LDS beside the exchange buffers (2^11 words at most, 16 KiB).
#ifndef SHG_KS_STAGE0_TABLE
#define SHG_KS_STAGE0_TABLE 1
#endif
constexpr int kKsTabLogB = 11;

// keyswitch_n14_kernel: its first in-wave exchange through cross-lane
// operations (v_permlane32_swap, v_permlane16_swap, DPP) instead of LDS.
#ifndef SHG_KS14_X1_LANES
#define SHG_KS14_X1_LANES 0
#endif
// Ring degrees (bit log n) that take the wave-exchange kernel.
#ifndef SHG_KS_WAVE_LOGN
#define SHG_KS_WAVE_LOGN (1 << 12)
#endif


namespace shg {

// FP64 schedule for key switching when every digit is below q/64 (bases up
// to 2^43 for 50-bit moduli).  With the per-stage bound B -> 1.375 B + q/2
// (see F64Arith), six stages from digits below q/64 reach 7.8 q, so the
// first reduction waits until after stage 5 and the second follows five
// stages later (7.7 q).  When at most three stages follow the last reduction
// the transform output stays below 3.5 q, each key product below
// q/2 + 0.375 * 3.5 q < 1.85 q, and three of them added to a reduced or
// canonical accumulator (<= q) stay below 8 q: the output is then used
// unreduced and the accumulator is reduced every three digits.
struct F64KsArith : F64Arith {
  using F64Arith::F64Arith;
  // Twiddles through 64-bit global addresses: buffer loads (as in F64Arith)
  // raised this kernel's VGPRs from 226 to 242 and measured 1.8% slower.
  __device__ __forceinline__ TwRaw fwd_tw_raw(int i) const { return fw[i]; }
  __device__ __forceinline__ Tw fwd_tw(int i) const { return tw(fw[i]); }
  // Without prep (small n) the accumulation relies on the last reduction.
  static constexpr bool kSkipLastReduce = false;
  __device__ static constexpr bool fwd_reduce_at(int s) { return s == 5 || s == 10; }
  __host__ __device__ static constexpr int last_reduce(int logn) {
    return logn - 1 >= 10 ? 10 : (logn - 1 >= 5 ? 5 : -1);
  }
  __host__ __device__ static constexpr bool ks_prep(int logn) {
    return logn - 1 - last_reduce(logn) > 3;
  }
  __host__ __device__ static constexpr int ks_acc_every(int logn) { return ks_prep(logn) ? 8 : 3; }
};

// Keys held as balanced FP64 values (shg_ks_key_create, |w| <= q/2): the
// multiply-accumulate takes each key word as it is, where F64KsArith
// converts it (an OR and a subtraction per word), and the tighter bounds
// allow a lazier schedule.  With |w| <= q/2 (balanced twiddles and keys) an
// FP64 product has |t| <= q/2 + 3 * 2^-53 |x| |w| <= q/2 + 0.1875 |x|, so a
// Cooley-Tukey stage maps a bound B to 1.1875 B + q/2.  Stage 0 (digits
// below 2^16 <= q/64) leaves B <= 0.532 q; stages 1-7 reach 7.99 q < 8 q,
// one reduction follows stage 7 (B <= q/2), and stages 8-13 reach 6.22 q.
// Those outputs enter the multiply-accumulate unreduced: each term is below
// q/2 + 0.1875 * 6.22 q = 1.67 q, so four of them on a canonical or reduced
// accumulator (<= q) stay below 7.7 q < 8 q < 2^53, and every sum and
// product term is an exact integer.  Same residues, canonical outputs.
struct F64KsKeyArith : F64KsArith {
  using F64KsArith::F64KsArith;
  __device__ __forceinline__ void mac(Acc& a, Prep d, uint64_t key, const PrimeDev&) const {
    a += f64_mulmod_h(d, __longlong_as_double((long long)key), q, qinv);
  }
  __device__ static constexpr bool fwd_reduce_at(int s) { return s == 7; }
  __host__ __device__ static constexpr int last_reduce(int logn) { return logn - 1 >= 7 ? 7 : -1; }
  __host__ __device__ static constexpr bool ks_prep(int) { return false; }
  __host__ __device__ static constexpr int ks_acc_every(int) { return 4; }
};

// LOGN = LOGM + HB.  With HB = 1 (n = 2^14) the limb is split between two
// workgroups: each loads both halves of the digit, applies the first
// Cooley-Tukey stage for its own half only, and runs the remaining stages on
// a 2^13-word slice, which keeps the accumulators of 16 words per thread in
// VGPRs without spilling.
//
// Workgroup order: logical index (l * count + ct) * 2^HB + h, limb-major.
// Hardware hands workgroup b to XCD b mod 8; the logical index is remapped
// so that each XCD runs one contiguous range (one output limb at the bench
// sizes): its CUs then share that limb's key rows and both halves of a
// ciphertext's source in their own L2.
template <int LOGM, int LOGE, int HB, class A, bool PL>
__global__ void __launch_bounds__((1 << LOGM) >> LOGE)
    keyswitch_kernel(KeySwitchArgs ka, GadgetDesc gd, const PrimeDev* __restrict__ primes,
                     LimbPrimes lp, int count, int num_limbs) {
  using V = typename A::V;
  using Acc = typename A::Acc;
  constexpr int E = 1 << LOGE;
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int M = 1 << LOGM;
  constexpr int LOGN = LOGM + HB;
  constexpr bool kPrep = A::ks_prep(LOGN);
  constexpr int kAccEvery = A::ks_acc_every(LOGN);
  // Double-buffered LDS exchanges (two slices; one workgroup per CU is
  // resident anyway, its VGPRs limit it).
  using Passes = FwdPasses<LOGM, LOGE, 0, A, false, HB, SHG_KS_DB != 0>;
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds0 = reinterpret_cast<V*>(smem);
  V* lds1 = lds0 + lds_words<LOGM>();
  unsigned b = blockIdx.x;
#if SHG_KS_XCD_REMAP
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);
#endif
  const int h = HB ? (int)(b & ((1u << HB) - 1)) : 0;
  const int lc = (int)(b >> HB);  // l * count + ct: limb-major order
  const int l = lc / count;
  const int ct = lc - l * count;
  const int t = threadIdx.x;
  const PrimeDev& P = primes[lp.id[l]];
  const A ar(P);
  const size_t limb_off = ((size_t)l << LOGN) + ((size_t)h << LOGM);
  const size_t poly_words = (size_t)num_limbs << LOGN;

  // Thread t accumulates the words the last register pass leaves it:
  // for each group gamma = gg*T + t, the 2^R consecutive positions
  // (gamma << R) + e of the slice (FwdLast); they are read and written with
  // 16-byte accesses.
  constexpr int R = FwdLast<LOGM, LOGE>::R_LAST;
  constexpr int G = FwdLast<LOGM, LOGE>::G_LAST;
  static_assert(R >= 1, "key switching needs at least one register stage");
  Acc acc0[E], acc1[E];
  uint64_t* o0 = ka.out0 + ct * ka.out_stride + limb_off;
  uint64_t* o1 = ka.out1 + ct * ka.out_stride + limb_off;
  {
    const uint64_t* i0 = ka.accumulate ? o0 : (ka.c0 ? ka.c0 + ct * ka.c0_stride + limb_off : nullptr);
    const uint64_t* i1 = ka.accumulate ? o1 : (ka.c1 ? ka.c1 + ct * ka.c1_stride + limb_off : nullptr);
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
      const size_t base = (size_t)(gg * T + t) << R;
#pragma unroll
      for (int e = 0; e < (1 << R); e += 2) {
        u64x2 a = i0 ? *reinterpret_cast<const u64x2*>(i0 + base + e) : u64x2{0, 0};
        u64x2 b2 = i1 ? *reinterpret_cast<const u64x2*>(i1 + base + e) : u64x2{0, 0};
        acc0[gg * (1 << R) + e] = ar.acc_in(a.x);
        acc0[gg * (1 << R) + e + 1] = ar.acc_in(a.y);
        acc1[gg * (1 << R) + e] = ar.acc_in(b2.x);
        acc1[gg * (1 << R) + e + 1] = ar.acc_in(b2.y);
      }
    }
  }

  // Forward transform of digit j's x (stage 0 already applied for HB = 1)
  // and its multiply-accumulate with key row j.
  // `issue` runs after the first key words are requested: the transform's
  // twiddle loads have been waited for by then, and memory loads complete
  // in order, so loads issued there (the next digit's plane words) hold up
  // neither the transform nor the first products.
  auto transform_and_mac = [&](int j, V (&x)[E], auto&& issue) {
    Passes::run(x, lds0, lds1, t, h, ar);
    if constexpr (SHG_KS_DB && (Passes::kExchanges & 1)) {
      V* tmp = lds0;
      lds0 = lds1;
      lds1 = tmp;
    }
    const uint64_t* kb = ka.key_b + (size_t)(ka.key_base + j) * poly_words + limb_off;
    const uint64_t* kap = ka.key_a + (size_t)(ka.key_base + j) * poly_words + limb_off;
#if SHG_KS_KEY_BUF
    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(kb), (short)0, 0x7fffffff, 0x00020000);
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(kap), (short)0, 0x7fffffff, 0x00020000);
#endif
#if SHG_KS_KEY_BUF
    // The key words are requested SHG_KS_KEY_BATCH 16-byte pairs (of each
    // key polynomial) at a time ahead of their products, so the load
    // latency is paid once per batch rather than once per pair.
    constexpr int kPairs = E / 2;
    constexpr int kBatch = SHG_KS_KEY_BATCH < kPairs ? SHG_KS_KEY_BATCH : kPairs;
    static_assert(kPairs % kBatch == 0, "key batches must tile the pairs");
#pragma unroll
    for (int p0 = 0; p0 < kPairs; p0 += kBatch) {
      U64Pair kbr[kBatch], kar[kBatch];
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int r = 2 * (p0 + i);
        const int gg = r >> R, e = r & ((1 << R) - 1);
        const int off = (int)((((uint32_t)(gg * T + t) << R) + e) * 8u);
        kbr[i] = buffer_load_u64x2(rb, off, 0);
        kar[i] = buffer_load_u64x2(ra, off, 0);
      }
      if (p0 == 0) issue();
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int r = 2 * (p0 + i);
        typename A::Prep d0, d1;
        if constexpr (kPrep) {
          d0 = ar.prep(x[r]);
          d1 = ar.prep(x[r + 1]);
        } else {
          d0 = x[r];
          d1 = x[r + 1];
        }
        ar.mac(acc0[r], d0, kbr[i].x, P);
        ar.mac(acc0[r + 1], d1, kbr[i].y, P);
        ar.mac(acc1[r], d0, kar[i].x, P);
        ar.mac(acc1[r + 1], d1, kar[i].y, P);
      }
    }
#else
    issue();
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
      const size_t base = (size_t)(gg * T + t) << R;
#pragma unroll
      for (int e = 0; e < (1 << R); e += 2) {
        const u64x2 kbv = *reinterpret_cast<const u64x2*>(kb + base + e);
        const u64x2 kav = *reinterpret_cast<const u64x2*>(kap + base + e);
        const int r = gg * (1 << R) + e;
        typename A::Prep d0, d1;
        if constexpr (kPrep) {
          d0 = ar.prep(x[r]);
          d1 = ar.prep(x[r + 1]);
        } else {
          d0 = x[r];
          d1 = x[r + 1];
        }
        ar.mac(acc0[r], d0, kbv.x, P);
        ar.mac(acc0[r + 1], d1, kbv.y, P);
        ar.mac(acc1[r], d0, kav.x, P);
        ar.mac(acc1[r + 1], d1, kav.y, P);
      }
    }
#endif
    if (((j + 1) % kAccEvery) == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        ar.acc_reduce(acc0[e]);
        ar.acc_reduce(acc1[e]);
      }
    }
    // The buffer the next digit writes first was last read before the
    // barrier of this digit's final exchange, so no barrier is needed.
  };

  if constexpr (PL) {
    // Digits from 16-bit planes: thread t's VC digits of a plane are
    // contiguous (VC / 8 16-byte loads); the next digit's are requested
    // after this one's transform, behind its first key words.
    constexpr int VC = ks_planes_vc(LOGN);
    // HB = 1 with digits of at most kKsTabLogB bits: the stage-0 products
    // w * v for every digit value v (the very values ct_side forms, so the
    // transform sees the same words), built once per workgroup and ordered
    // before the first digit by a barrier of their own.
    constexpr bool kTab = HB == 1 && SHG_KS_STAGE0_TABLE && std::is_same<V, double>::value;
    const bool use_tab = kTab && gd.max_log_b <= kKsTabLogB;
    V* tab = reinterpret_cast<V*>(smem) + lds_words<LOGM>() * (SHG_KS_DB ? 2 : 1);
    if constexpr (kTab) {
      if (use_tab) {
        const auto w = ar.fwd_tw(1);
        for (int v = t; v < (1 << gd.max_log_b); v += T) tab[v] = ar.ct_product((double)v, w);
        __syncthreads();
      }
    }
    constexpr int NW = VC / 8;
    static_assert(VC == E << HB, "plane layout must match the slice tiling");
    const uint4* pl = reinterpret_cast<const uint4*>(ka.planes + ((size_t)ct * gd.k << LOGN)) +
                      (size_t)t * NW;
    constexpr size_t kPlane = ((size_t)1 << LOGN) / 8;  // uint4 per plane
    uint4 cur[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) cur[w] = pl[w];
    for (int j = 0; j < gd.k; ++j) {
      uint4 nxt[NW];
      uint32_t wd[VC / 2];
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        wd[4 * w] = cur[w].x;
        wd[4 * w + 1] = cur[w].y;
        wd[4 * w + 2] = cur[w].z;
        wd[4 * w + 3] = cur[w].w;
      }
      auto digit = [&](int i) {
        const uint32_t v = (i & 1) ? (wd[i / 2] >> 16) : (wd[i / 2] & 0xffffu);
        return ar.digit_in32(v, P);
      };
      V x[E];
      if constexpr (HB == 0) {
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = digit(e);
      } else {
        // Stage 0 for this half: u + w v (h = 0) or u - w v (h = 1).
        const auto w = ar.fwd_tw(1);
        const double sgn = h ? -1.0 : 1.0;
        static_assert(!A::fwd_reduce_at(0), "stage 0 is not followed by a reduction here");
        bool done = false;
        if constexpr (kTab) {
          if (use_tab) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
              const int i = E + e;
              const uint32_t v = (i & 1) ? (wd[i / 2] >> 16) : (wd[i / 2] & 0xffffu);
              x[e] = fma(sgn, tab[v], digit(e));
            }
            done = true;
          }
        }
        if (!done) {
#pragma unroll
          for (int e = 0; e < E; ++e) x[e] = ar.ct_side(digit(e), digit(E + e), w, sgn);
        }
      }
      // The last digit reloads its own plane (unconditional loads keep the
      // compiler's count of outstanding loads exact at every wait).
      const size_t jn = j + 1 < gd.k ? (size_t)(j + 1) : (size_t)j;
      transform_and_mac(j, x, [&] {
#pragma unroll
        for (int w = 0; w < NW; ++w) nxt[w] = pl[jn * kPlane + w];
      });
#pragma unroll
      for (int w = 0; w < NW; ++w) cur[w] = nxt[w];
    }
  } else {
    const uint64_t* src_ct = ka.src + ct * ka.src_stride;
    const bool digits32 = gd.max_log_b <= 32;
    int j = 0;
    // Digits of source limb i are consecutive: j runs over (i, dd).
    for (int i = 0; i < num_limbs; ++i) {
      const uint64_t* src = src_ct + ((size_t)i << LOGN);
      const int lb = gd.limb_log_b[i];
      const int nd = gd.limb_digits[i];
      const uint64_t mask = (1ull << lb) - 1;
      for (int dd = 0; dd < nd; ++dd, ++j) {
        const int sh = dd * lb;
        V x[E];
        // Digit extraction (one uniform branch per digit on the base
        // width), then for HB = 1 the first Cooley-Tukey stage on this half.
        auto load_digits = [&](auto conv) {
          if constexpr (HB == 0) {
#pragma unroll
            for (int e = 0; e < E; ++e) x[e] = conv(src[e * T + t] >> sh);
          } else {
            const auto w = ar.fwd_tw(1);
#pragma unroll
            for (int e = 0; e < E; ++e) {
              V u = conv(src[e * T + t] >> sh);
              V v = conv(src[M + e * T + t] >> sh);
              ar.ct(u, v, w);
              x[e] = h ? v : u;
              if (ar.fwd_reduce_at(0)) ar.fwd_reduce(x[e]);
            }
          }
        };
        if (gd.limb_balance[i]) {
          const uint64_t qs = primes[lp.id[i]].q;
          load_digits([&](uint64_t s64) { return ar.digit_in_bal(s64 & mask, qs, P); });
        } else if (digits32) {
          const uint32_t m32 = (uint32_t)mask;
          load_digits([&](uint64_t s64) { return ar.digit_in32((uint32_t)s64 & m32, P); });
        } else {
          load_digits([&](uint64_t s64) { return ar.digit_in(s64 & mask, P); });
        }
        transform_and_mac(j, x, [] {});
      }
    }
  }
  // The thread's output offsets are formed again here from an opaque copy
  // of its index: kept from the initial loads across the digit loop they
  // would be spilled (at 8 words per thread the loop needs all 128 VGPRs).
  int t_out = threadIdx.x;
  asm volatile("" : "+v"(t_out));
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const size_t base = (size_t)(gg * T + t_out) << R;
#pragma unroll
    for (int e = 0; e < (1 << R); e += 2) {
      u64x2 a, b2;
      a.x = ar.acc_out(acc0[gg * (1 << R) + e]);
      a.y = ar.acc_out(acc0[gg * (1 << R) + e + 1]);
      b2.x = ar.acc_out(acc1[gg * (1 << R) + e]);
      b2.y = ar.acc_out(acc1[gg * (1 << R) + e + 1]);
      *reinterpret_cast<u64x2*>(o0 + base + e) = a;
      *reinterpret_cast<u64x2*>(o1 + base + e) = b2;
    }
  }
}

// The FP64 digit-plane key switch at n = 2^14 with one workgroup-wide
// exchange per digit.  A workgroup of 16 waves x 64 lanes x 8 words takes
// half h of output limb l of one ciphertext (positions P = 2^13 h + p of the
// limb); stage 0 combines both halves of the digit, stages 1-13 run on the
// slice p (13 bits) through five layouts of word e of thread (wave w, lane l):
//   A: p = w + 16 l + 2^10 e                      stages 1-3 (bits 12..10)
//   B: p = w + 16 (l & 7) + 2^7 e + 2^10 (l >> 3) stages 4-6 (bits 9..7)
//   C: p = w + 16 e + 2^7 (l & 7) + 2^10 (l >> 3) stages 7-9 (bits 6..4)
//   D: p = b0 + 2 e + 16 (l & 15) + 2^8 b8 + 2^9 w, b0 = l bit 4, b8 = l bit 5
//                                                 stages 10-12 (bits 3..1)
//   E: p = (e & 1) + 2 l + 2^7 (e >> 1) + 2^9 w   stage 13 (bit 0), products
// A -> B, B -> C and D -> E stay inside each wave (private LDS regions of 575
// words; no barrier), C -> D crosses waves (one barrier).  Stage s of the
// limb transform takes psi_bitrev[2^s + (P >> (14 - s))]: per pass of three
// stages a run of 1, 2 and 4 twiddles from a base B0 (h in pass 1, through
// the scalar cache; 8 h + (l >> 3), 64 h + l and 2^9 h + 2^5 w + 16 b8 +
// (l & 15) after), and psi_bitrev[2^13 + 2^12 h + 2^8 w + 64 k + l] for the
// pair k of stage 13.  Exchange C -> D has a single buffer and no second
// barrier: its slot of p is 546 r + 34 a + m + (m >> 4) with m = p bits
// 8..4 and (r, a) = (p >> 9, p & 15) for even digits, (p & 15, p >> 9) for
// odd ones, so each wave writes a digit's words exactly where it read the
// previous digit's (r names the reading wave in even digits and the writing
// wave in odd ones) and no wave overwrites words another has yet to read.
// Every slot map keeps the register index out of the lane part (immediate
// LDS offsets) and no two lanes of a 16-lane group share a bank pair.  The
// digit planes hold, in row t = 64 w + l, the 16 digits of limb positions
// w + 16 l + 2^10 k (k < 8: half 0), so thread t reads row t (16-bit entries,
// or packed at the common base log: plane_pack_bits).
namespace ks14 {
constexpr int kRegion = 575;                  // in-wave exchange words per wave
constexpr int kX3Words = 8733;                // workgroup exchange words
constexpr int kLdsWords = kX3Words + 16 * kRegion;  // then the stage-0 table
}  // namespace ks14

// Layout A -> B without LDS: B swaps register bit k with lane bit 3 + k
// (k = 0, 1, 2).  Each is a transposition of a register pair (x[e], x[e |
// 2^k], bit k of e clear) with a lane bit m: lanes with bit m clear take
// the partner lane's x[e] into x[e | 2^k], lanes with it set the partner's
// x[e | 2^k] into x[e].  Lane bit 5: v_permlane32_swap (lanes 32-63 of the
// first operand with lanes 0-31 of the second); bit 4: v_permlane16_swap
// (odd rows of 16 with even rows); bit 3: DPP row_ror:8 (lane l ^ 8 of the
// row) written to banks 2-3 or 0-1 of each row.
__device__ __forceinline__ void ks14_lane_transpose(double (&x)[8]) {
  auto split = [](double v, uint32_t& lo, uint32_t& hi) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    lo = (uint32_t)u;
    hi = (uint32_t)(u >> 32);
  };
  auto join = [](uint32_t lo, uint32_t hi) {
    return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
  };
#pragma unroll
  for (int e = 0; e < 4; ++e) {  // register bit 2 <-> lane bit 5
    uint32_t al, ah, bl, bh;
    split(x[e], al, ah);
    split(x[e | 4], bl, bh);
    const auto lo = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
    x[e] = join(lo[0], hi[0]);
    x[e | 4] = join(lo[1], hi[1]);
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {  // register bit 1 <-> lane bit 4
    if (e & 2) continue;
    uint32_t al, ah, bl, bh;
    split(x[e], al, ah);
    split(x[e | 2], bl, bh);
    const auto lo = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
    x[e] = join(lo[0], hi[0]);
    x[e | 2] = join(lo[1], hi[1]);
  }
#pragma unroll
  for (int e = 0; e < 8; e += 2) {  // register bit 0 <-> lane bit 3
    uint32_t al, ah, bl, bh;
    split(x[e], al, ah);
    split(x[e | 1], bl, bh);
    const uint32_t nal = __builtin_amdgcn_update_dpp(al, bl, 0x128, 0xf, 0xc, false);
    const uint32_t nah = __builtin_amdgcn_update_dpp(ah, bh, 0x128, 0xf, 0xc, false);
    const uint32_t nbl = __builtin_amdgcn_update_dpp(bl, al, 0x128, 0xf, 0x3, false);
    const uint32_t nbh = __builtin_amdgcn_update_dpp(bh, ah, 0x128, 0xf, 0x3, false);
    x[e] = join(nal, nah);
    x[e | 1] = join(nbl, nbh);
  }
}

template <class A, int S0>
__device__ __forceinline__ void ks14_pass(double (&x)[8], const double (&tw)[7], const A& ar) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int kb = 2 - j;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if (e & (1 << kb)) continue;
      ar.ct(x[e], x[e | (1 << kb)], ar.tw(tw[(1 << j) - 1 + (e >> (kb + 1))]));
    }
    if (A::fwd_reduce_at(S0 + j) && !(A::kSkipLastReduce && S0 + j == 13)) {
#pragma unroll
      for (int e = 0; e < 8; ++e) ar.fwd_reduce(x[e]);
    }
  }
}

// Twiddles of the pass starting at stage S0 from base b0: psi_bitrev[2^S0 +
// b0], [2^(S0+1) + 2 b0 + 0..1], [2^(S0+2) + 4 b0 + 0..3].
template <int S0>
__device__ __forceinline__ void ks14_tw(__amdgpu_buffer_rsrc_t r, int b0, double (&tw)[7]) {
  w32::tw_run<1>(r, (1 << S0) + b0, tw);
  w32::tw_run<2>(r, (2 << S0) + 2 * b0, tw + 1);
  w32::tw_run<4>(r, (4 << S0) + 4 * b0, tw + 3);
}

template <class A, int PB>
__global__ void __launch_bounds__(1024)
    keyswitch_n14_kernel(KeySwitchArgs ka, GadgetDesc gd, const PrimeDev* __restrict__ primes,
                         LimbPrimes lp, int count, int num_limbs) {
  using cptr = const double __attribute__((address_space(4)))*;
  constexpr int E = 8;
  constexpr int LOGN = 14;
  constexpr bool kPrep = A::ks_prep(LOGN);
  constexpr int kAccEvery = A::ks_acc_every(LOGN);
  extern __shared__ __align__(16) unsigned char smem[];
  double* x3 = reinterpret_cast<double*>(smem);
  unsigned b = blockIdx.x;
#if SHG_KS_XCD_REMAP
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);
#endif
  const int h = (int)(b & 1);
  const int lc = (int)(b >> 1);  // l * count + ct: limb-major order
  const int l = lc / count;
  const int ct = lc - l * count;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const PrimeDev& P = primes[lp.id[l]];
  const A ar(P);
  const __amdgpu_buffer_rsrc_t twr = w32::rsrc(P.fwd);
  const cptr ctw = (cptr)(size_t)(const void*)P.fwd;
  double* rg = x3 + ks14::kX3Words + wave * ks14::kRegion;  // this wave's region
  const size_t limb_off = ((size_t)l << LOGN) + ((size_t)h << (LOGN - 1));
  const size_t poly_words = (size_t)num_limbs << LOGN;

  // Layout E word offsets: pair k = e >> 1 at 2 l + 128 k + 512 w.
  const int pe = 2 * lane + 512 * wave;
  double acc0[E], acc1[E];
  {
    const uint64_t* i0 = ka.accumulate ? ka.out0 + ct * ka.out_stride
                                       : (ka.c0 ? ka.c0 + ct * ka.c0_stride : nullptr);
    const uint64_t* i1 = ka.accumulate ? ka.out1 + ct * ka.out_stride
                                       : (ka.c1 ? ka.c1 + ct * ka.c1_stride : nullptr);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t o = limb_off + pe + 128 * k;
      u64x2 a = i0 ? *reinterpret_cast<const u64x2*>(i0 + o) : u64x2{0, 0};
      u64x2 c = i1 ? *reinterpret_cast<const u64x2*>(i1 + o) : u64x2{0, 0};
      acc0[2 * k] = ar.acc_in(a.x);
      acc0[2 * k + 1] = ar.acc_in(a.y);
      acc1[2 * k] = ar.acc_in(c.x);
      acc1[2 * k + 1] = ar.acc_in(c.y);
    }
  }

  // Stage-0 products w v for every digit value v (digits of at most
  // kKsTabLogB bits), as in keyswitch_kernel.
  const bool use_tab = gd.max_log_b <= kKsTabLogB;
  double* tab = x3 + ks14::kLdsWords;
  if (use_tab) {
    const auto w1 = ar.fwd_tw(1);
    for (int v = t; v < (1 << gd.max_log_b); v += 1024) tab[v] = ar.ct_product((double)v, w1);
    __syncthreads();
  }
  const auto w1 = ar.fwd_tw(1);
  const double sgn = h ? -1.0 : 1.0;

  // Lane parts of the exchange slots (the register part is an immediate).
  const int b0 = (lane >> 4) & 1, b8 = (lane >> 5) & 1;
  const int x1w = lane;                                   // + 72 e
  const int x1r = (lane & 7) + 72 * (lane >> 3);          // + 8 e
  const int x2w = x1r;                                    // + 9 e
  const int x2r = 9 * (lane & 7) + 72 * (lane >> 3);      // + e
  const int x3w0 = (lane >> 2) * 546 + 34 * wave + 8 * (lane & 3) + ((lane & 3) >> 1);  // + e
  const int x3w1 = 546 * wave + 34 * (lane >> 2) + 8 * (lane & 3) + ((lane & 3) >> 1);  // + e
  const int x3r0 = 546 * wave + 34 * b0 + (lane & 15) + 17 * b8;                         // + 68 e
  const int x3r1 = 546 * b0 + 34 * wave + (lane & 15) + 17 * b8;                         // + 1092 e
  const int x4w = b0 + 17 * (lane & 15) + 272 * b8;       // + 2 e
  const int x4r = 2 * lane + (lane >> 3);                 // + (e & 1) + 136 (e >> 1)
  // Twiddle bases of passes 2-5.
  const int tb2 = 8 * h + (lane >> 3);
  const int tb3 = 64 * h + lane;
  const int tb4 = 512 * h + 32 * wave + 16 * b8 + (lane & 15);
  const int tb5 = 8192 + 4096 * h + 256 * wave + lane;

  // Digit planes: row t holds this thread's 16 digits of the plane, as 16-bit
  // entries (PB = 0: two 16-byte loads) or packed at PB bits (PW dwords, the
  // k-th at dword k * 1024 + t of the plane).
  constexpr int PW = PB ? (16 * PB + 31) / 32 : 8;
  const uint32_t* pl = reinterpret_cast<const uint32_t*>(ka.planes + ((size_t)ct * gd.k << LOGN));
  constexpr size_t kPlane = ((size_t)1 << LOGN) / 2;  // dwords per plane
  auto load_row = [&](size_t j, uint32_t (&w)[8]) {
    if constexpr (PB == 0) {
      const uint4* r = reinterpret_cast<const uint4*>(pl + j * kPlane) + (size_t)t * 2;
      const uint4 a = r[0], c = r[1];
      w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
      w[4] = c.x; w[5] = c.y; w[6] = c.z; w[7] = c.w;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) w[q] = q < PW ? pl[j * kPlane + q * 1024 + t] : 0u;
    }
  };
  uint32_t cur[8];
  load_row(0, cur);
  for (int j = 0; j < gd.k; ++j) {
    double x[E];
    {
      uint32_t dv[16];
      if constexpr (PB == 0) {
#pragma unroll
        for (int i = 0; i < 16; ++i) dv[i] = (i & 1) ? (cur[i / 2] >> 16) : (cur[i / 2] & 0xffffu);
      } else {
        unpack_row16<PB>(cur, dv);
      }
      auto dig = [&](int i) { return dv[i]; };
      if (use_tab) {
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = fma(sgn, tab[dig(E + e)], ar.digit_in32(dig(e), P));
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e)
          x[e] = ar.ct_side(ar.digit_in32(dig(e), P), ar.digit_in32(dig(E + e), P), w1, sgn);
      }
    }
    double tw[7];
    // Pass 1 (layout A), uniform twiddles.
    tw[0] = ctw[2 + h];
    tw[1] = ctw[4 + 2 * h];
    tw[2] = ctw[5 + 2 * h];
#pragma unroll
    for (int c = 0; c < 4; ++c) tw[3 + c] = ctw[8 + 4 * h + c];
    ks14_pass<A, 1>(x, tw, ar);
    // A -> B inside the wave.
    ks14_tw<4>(twr, tb2, tw);
    if constexpr (SHG_KS14_X1_LANES) {
      ks14_lane_transpose(x);
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) rg[x1w + 72 * e] = x[e];
      wave_sync_lds();
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = rg[x1r + 8 * e];
      pin_words(x);
      wave_sync_lds();
    }
    ks14_pass<A, 4>(x, tw, ar);
    // B -> C inside the wave.
    ks14_tw<7>(twr, tb3, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) rg[x2w + 9 * e] = x[e];
    wave_sync_lds();
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = rg[x2r + e];
    pin_words(x);
    wave_sync_lds();
    ks14_pass<A, 7>(x, tw, ar);
    // C -> D across the workgroup.
    ks14_tw<10>(twr, tb4, tw);
    if (j & 1) {
#pragma unroll
      for (int e = 0; e < E; ++e) x3[x3w1 + e] = x[e];
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = x3[x3r1 + 1092 * e];
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) x3[x3w0 + e] = x[e];
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = x3[x3r0 + 68 * e];
    }
    pin_words(x);
    ks14_pass<A, 10>(x, tw, ar);
    // D -> E inside the wave, then stage 13 on register bit 0.
    double t5[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) t5[k] = __builtin_bit_cast(double, w32::ld64(twr, (tb5 + 64 * k) * 8, 0));
#pragma unroll
    for (int e = 0; e < E; ++e) rg[x4w + 2 * e] = x[e];
    wave_sync_lds();
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = rg[x4r + (e & 1) + 136 * (e >> 1)];
    pin_words(x);
    wave_sync_lds();
#pragma unroll
    for (int k = 0; k < 4; ++k) ar.ct(x[2 * k], x[2 * k + 1], ar.tw(t5[k]));
    if (A::fwd_reduce_at(13) && !A::kSkipLastReduce) {
#pragma unroll
      for (int e = 0; e < E; ++e) ar.fwd_reduce(x[e]);
    }
    // Products with key row j: both key polynomials' pairs requested at
    // once, then the next digit's plane words behind them.
    const uint64_t* kbp = ka.key_b + (size_t)(ka.key_base + j) * poly_words + limb_off;
    const uint64_t* kap = ka.key_a + (size_t)(ka.key_base + j) * poly_words + limb_off;
    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(kbp), (short)0, 0x7fffffff, 0x00020000);
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(kap), (short)0, 0x7fffffff, 0x00020000);
    const size_t jn = j + 1 < gd.k ? (size_t)(j + 1) : (size_t)j;
    uint32_t nxt[8];
#pragma unroll
    for (int k0 = 0; k0 < 4; k0 += 2) {
      U64Pair kbr[2], kar[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        kbr[i] = buffer_load_u64x2(rb, (pe + 128 * (k0 + i)) * 8, 0);
        kar[i] = buffer_load_u64x2(ra, (pe + 128 * (k0 + i)) * 8, 0);
      }
      if (k0 == 0) load_row(jn, nxt);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = 2 * (k0 + i);
        typename A::Prep d0, d1;
        if constexpr (kPrep) {
          d0 = ar.prep(x[r]);
          d1 = ar.prep(x[r + 1]);
        } else {
          d0 = x[r];
          d1 = x[r + 1];
        }
        ar.mac(acc0[r], d0, kbr[i].x, P);
        ar.mac(acc0[r + 1], d1, kbr[i].y, P);
        ar.mac(acc1[r], d0, kar[i].x, P);
        ar.mac(acc1[r + 1], d1, kar[i].y, P);
      }
    }
    if (((j + 1) % kAccEvery) == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        ar.acc_reduce(acc0[e]);
        ar.acc_reduce(acc1[e]);
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) cur[q] = nxt[q];
  }
  // Output offsets formed again from an opaque copy of the thread index.
  int t_out = threadIdx.x;
  asm volatile("" : "+v"(t_out));
  const int po = 2 * (t_out & 63) + 512 * (t_out >> 6);
  uint64_t* o0 = ka.out0 + ct * ka.out_stride + limb_off;
  uint64_t* o1 = ka.out1 + ct * ka.out_stride + limb_off;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    u64x2 a, c;
    a.x = ar.acc_out(acc0[2 * k]);
    a.y = ar.acc_out(acc0[2 * k + 1]);
    c.x = ar.acc_out(acc1[2 * k]);
    c.y = ar.acc_out(acc1[2 * k + 1]);
    *reinterpret_cast<u64x2*>(o0 + po + 128 * k) = a;
    *reinterpret_cast<u64x2*>(o1 + po + 128 * k) = c;
  }
}

template <class A>
static hipError_t launch_ks_n14(const KeySwitchArgs& ka, const GadgetDesc& gd,
                                const PrimeDev* primes, const LimbPrimes& lp, int count,
                                int num_limbs, hipStream_t s) {
  const bool tab = gd.max_log_b <= kKsTabLogB;
  const size_t shmem = (size_t)ks14::kLdsWords * 8 + (tab ? ((size_t)8 << gd.max_log_b) : 0);
  const int pb = plane_pack_bits(gd.plane_bits);
  auto k = pb == 8 ? keyswitch_n14_kernel<A, 8>
           : pb == 10 ? keyswitch_n14_kernel<A, 10>
           : pb == 12 ? keyswitch_n14_kernel<A, 12> : keyswitch_n14_kernel<A, 0>;
  set_max_dynamic_shmem((const void*)k, shmem);
  hipLaunchKernelGGL(k, dim3((unsigned)(count * num_limbs) << 1), dim3(1024), shmem, s, ka, gd,
                     primes, lp, count, num_limbs);
  return hipGetLastError();
}

// The digit-plane key switch on the wave-exchange layout (shg_ntt_wave.h):
// slices of 2^(10+LW) words (n = 2^12: LW = 2; 2^13: LW = 3; 2^14: LW = 3
// with HB = 1), FP64.  Per digit: one workgroup-wide LDS exchange between
// two barriers, the others inside each wave; uniform twiddles through the
// scalar cache.  The multiply-accumulate runs in the last pass's layout:
// thread (wave w, lane l) holds, for gg = 0..3, the four consecutive words
// tp4 + 256 gg + 0..3 of the slice.  Workgroup order and remap as in
// keyswitch_kernel.
template <int LW, int HB, class A>
__global__ void __launch_bounds__(64 << LW)
    keyswitch_wave_kernel(KeySwitchArgs ka, GadgetDesc gd, const PrimeDev* __restrict__ primes,
                          LimbPrimes lp, int count, int num_limbs) {
  using V = double;
  using Lay = WaveLayout<LW>;
  constexpr int LOGM = Lay::LOGM;
  constexpr int E = 16;
  constexpr int T = 64 << LW;
  constexpr int LOGN = LOGM + HB;
  constexpr bool kPrep = A::ks_prep(LOGN);
  constexpr int kAccEvery = A::ks_acc_every(LOGN);
  using T1 = WavePassTw<LW, 1, 3, 0>;
  using T2 = WavePassTw<LW, 2, LW - 1, 0>;
  using T3 = WavePassTw<LW, 3, 3, 0>;
  using T4 = WavePassTw<LW, 4, 1, 0>;
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds = reinterpret_cast<V*>(smem);
  unsigned b = blockIdx.x;
#if SHG_KS_XCD_REMAP
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);
#endif
  const int h = HB ? (int)(b & ((1u << HB) - 1)) : 0;
  const int lc = (int)(b >> HB);  // l * count + ct: limb-major order
  const int l = lc / count;
  const int ct = lc - l * count;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const PrimeDev& P = primes[lp.id[l]];
  const A ar(P);
  const size_t limb_off = ((size_t)l << LOGN) + ((size_t)h << LOGM);
  const size_t poly_words = (size_t)num_limbs << LOGN;

  V acc0[E], acc1[E];
  uint64_t* o0 = ka.out0 + ct * ka.out_stride + limb_off;
  uint64_t* o1 = ka.out1 + ct * ka.out_stride + limb_off;
  {
    const int tp4 = Lay::thread_part(4, lane, wave);
    const uint64_t* i0 = ka.accumulate ? o0 : (ka.c0 ? ka.c0 + ct * ka.c0_stride + limb_off : nullptr);
    const uint64_t* i1 = ka.accumulate ? o1 : (ka.c1 ? ka.c1 + ct * ka.c1_stride + limb_off : nullptr);
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const size_t base = (size_t)tp4 + Lay::reg_part(4, gg * 4);
#pragma unroll
      for (int e = 0; e < 4; e += 2) {
        u64x2 a = i0 ? *reinterpret_cast<const u64x2*>(i0 + base + e) : u64x2{0, 0};
        u64x2 b2 = i1 ? *reinterpret_cast<const u64x2*>(i1 + base + e) : u64x2{0, 0};
        acc0[gg * 4 + e] = ar.acc_in(a.x);
        acc0[gg * 4 + e + 1] = ar.acc_in(a.y);
        acc1[gg * 4 + e] = ar.acc_in(b2.x);
        acc1[gg * 4 + e + 1] = ar.acc_in(b2.y);
      }
    }
  }

  // Digits from 16-bit planes (launch_ntt_inv_digits): thread t's VC digits
  // of a plane are positions t + T v, v < VC, i.e. pass 1's layout (both
  // halves of the limb for HB = 1); the next digit's are loaded while this
  // one is transformed.
  constexpr int VC = ks_planes_vc(LOGN);
  constexpr int NW = VC / 8;
  static_assert(VC == E << HB, "plane layout must match the slice tiling");
  const uint4* pl = reinterpret_cast<const uint4*>(ka.planes + ((size_t)ct * gd.k << LOGN)) +
                    (size_t)t * NW;
  constexpr size_t kPlane = ((size_t)1 << LOGN) / 8;  // uint4 per plane
  uint4 cur[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) cur[w] = pl[w];
  for (int j = 0; j < gd.k; ++j) {
    uint4 nxt[NW];
    if (j + 1 < gd.k) {
#pragma unroll
      for (int w = 0; w < NW; ++w) nxt[w] = pl[(size_t)(j + 1) * kPlane + w];
    }
    // An opaque copy of the lane index per digit keeps the per-thread
    // twiddle and LDS offsets inside the loop (hoisted they would stay live
    // across it beside the accumulators).
    int lane_o = lane;
    asm volatile("" : "+v"(lane_o));
    const int tp1 = Lay::thread_part(1, lane_o, wave);
    double tw1[T1::N];
    wave_pass_tw<LW, 1, 3, 0, true, HB>(tw1, 0, h, ar);
    uint32_t wd[VC / 2];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      wd[4 * w] = cur[w].x;
      wd[4 * w + 1] = cur[w].y;
      wd[4 * w + 2] = cur[w].z;
      wd[4 * w + 3] = cur[w].w;
    }
    auto digit = [&](int i) {
      const uint32_t v = (i & 1) ? (wd[i / 2] >> 16) : (wd[i / 2] & 0xffffu);
      return ar.digit_in32(v, P);
    };
    V x[E];
    if constexpr (HB == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = digit(e);
    } else {
      const auto w = ar.fwd_tw(1);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        V u = digit(e);
        V v = digit(E + e);
        ar.ct(u, v, w);
        x[e] = h ? v : u;
        if (ar.fwd_reduce_at(0)) ar.fwd_reduce(x[e]);
      }
    }
    wave_pass<LW, 1, 3, 0, HB>(x, tw1, ar);
    __builtin_amdgcn_sched_barrier(0);
    double tw2[T2::N];
    wave_pass_tw<LW, 2, LW - 1, 0, true, HB>(tw2, wave << 10, h, ar);
    // Every wave has finished the previous digit's in-wave exchanges.
    __syncthreads();
    wave_to_lds<LW, 1>(x, lds, tp1);
    __syncthreads();
    const int tp2 = Lay::thread_part(2, lane_o, wave);
    wave_from_lds<LW, 2>(x, lds, tp2);
    wave_pass<LW, 2, LW - 1, 0, HB>(x, tw2, ar);
    const int tp3 = Lay::thread_part(3, lane_o, wave);
    __builtin_amdgcn_sched_barrier(0);
    double tw3[T3::N];
    wave_pass_tw<LW, 3, 3, 0, false, HB>(tw3, tp3, h, ar);
    wave_to_lds<LW, 2>(x, lds, tp2);
    wave_sync_lds();
    wave_from_lds<LW, 3>(x, lds, tp3);
    wave_pass<LW, 3, 3, 0, HB>(x, tw3, ar);
    const int tp4 = Lay::thread_part(4, lane_o, wave);
    __builtin_amdgcn_sched_barrier(0);
    double tw4[T4::N];
    wave_pass_tw<LW, 4, 1, 0, false, HB>(tw4, tp4, h, ar);
    wave_to_lds<LW, 3>(x, lds, tp3);
    wave_sync_lds();
    wave_from_lds<LW, 4>(x, lds, tp4);
    // The next digit's exchange writes these slots after its first barrier.
    pin_words(x);
    wave_pass<LW, 4, 1, 0, HB>(x, tw4, ar);
    // Multiply-accumulate with key row j.
#if SHG_KS_KEY_BUF
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint64_t*>(ka.key_b + (size_t)(ka.key_base + j) * poly_words + limb_off),
        (short)0, 0x7fffffff, 0x00020000);
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint64_t*>(ka.key_a + (size_t)(ka.key_base + j) * poly_words + limb_off),
        (short)0, 0x7fffffff, 0x00020000);
#else
    const uint64_t* kb = ka.key_b + (size_t)(ka.key_base + j) * poly_words + limb_off + tp4;
    const uint64_t* kap = ka.key_a + (size_t)(ka.key_base + j) * poly_words + limb_off + tp4;
#endif
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const int off = Lay::reg_part(4, gg * 4);
#pragma unroll
      for (int e = 0; e < 4; e += 2) {
#if SHG_KS_KEY_BUF
        const int bo = (int)((uint32_t)(tp4 + off + e) * 8u);
        const U64Pair kbv = buffer_load_u64x2(rb, bo, 0);
        const U64Pair kav = buffer_load_u64x2(ra, bo, 0);
#else
        const u64x2 kbv = *reinterpret_cast<const u64x2*>(kb + off + e);
        const u64x2 kav = *reinterpret_cast<const u64x2*>(kap + off + e);
#endif
        const int r = gg * 4 + e;
        typename A::Prep d0, d1;
        if constexpr (kPrep) {
          d0 = ar.prep(x[r]);
          d1 = ar.prep(x[r + 1]);
        } else {
          d0 = x[r];
          d1 = x[r + 1];
        }
        ar.mac(acc0[r], d0, kbv.x, P);
        ar.mac(acc0[r + 1], d1, kbv.y, P);
        ar.mac(acc1[r], d0, kav.x, P);
        ar.mac(acc1[r + 1], d1, kav.y, P);
      }
    }
    if (((j + 1) % kAccEvery) == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        ar.acc_reduce(acc0[e]);
        ar.acc_reduce(acc1[e]);
      }
    }
    if (j + 1 < gd.k) {
#pragma unroll
      for (int w = 0; w < NW; ++w) cur[w] = nxt[w];
    }
  }
  const int tp4 = Lay::thread_part(4, lane, wave);
#pragma unroll
  for (int gg = 0; gg < 4; ++gg) {
    const size_t base = (size_t)tp4 + Lay::reg_part(4, gg * 4);
#pragma unroll
    for (int e = 0; e < 4; e += 2) {
      u64x2 a, b2;
      a.x = ar.acc_out(acc0[gg * 4 + e]);
      a.y = ar.acc_out(acc0[gg * 4 + e + 1]);
      b2.x = ar.acc_out(acc1[gg * 4 + e]);
      b2.y = ar.acc_out(acc1[gg * 4 + e + 1]);
      *reinterpret_cast<u64x2*>(o0 + base + e) = a;
      *reinterpret_cast<u64x2*>(o1 + base + e) = b2;
    }
  }
}

template <int LW, int HB, class A = F64KsArith>
static hipError_t launch_ks_wave(const KeySwitchArgs& ka, const GadgetDesc& gd,
                                 const PrimeDev* primes, const LimbPrimes& lp, int count,
                                 int num_limbs, hipStream_t s) {
  constexpr int T = 64 << LW;
  const size_t shmem = (size_t)lds_words<10 + LW>() * 8;
  auto k = keyswitch_wave_kernel<LW, HB, A>;
  if (shmem > 65536) set_max_dynamic_shmem((const void*)k, shmem);
  hipLaunchKernelGGL(k, dim3((unsigned)(count * num_limbs) << HB), dim3(T), shmem, s, ka, gd,
                     primes, lp, count, num_limbs);
  return hipGetLastError();
}

template <int LOGM, int LOGE, class A, int HB, bool PL>
static hipError_t launch_ks(const KeySwitchArgs& ka, const GadgetDesc& gd, const PrimeDev* primes,
                            const LimbPrimes& lp, int count, int num_limbs, hipStream_t s) {
  constexpr int T = (1 << LOGM) >> LOGE;
  // Exchange buffers, then (split limbs with small digits) the stage-0 table.
  const bool tab = PL && HB == 1 && SHG_KS_STAGE0_TABLE && gd.max_log_b <= kKsTabLogB &&
                   std::is_same<typename A::V, double>::value;
  const size_t shmem = (size_t)lds_words<LOGM>() * 8 * (SHG_KS_DB ? 2 : 1) +
                       (tab ? ((size_t)8 << gd.max_log_b) : 0);
  auto k = keyswitch_kernel<LOGM, LOGE, HB, A, PL>;
  if (shmem > 65536) set_max_dynamic_shmem((const void*)k, shmem);
  hipLaunchKernelGGL(k, dim3((unsigned)(count * num_limbs) << HB), dim3(T), shmem, s, ka, gd,
                     primes, lp, count, num_limbs);
  return hipGetLastError();
}

template <class A, bool PL>
static hipError_t dispatch_ks(int logn, const KeySwitchArgs& ka, const GadgetDesc& gd,
                              const PrimeDev* primes, const LimbPrimes& lp, int count,
                              int num_limbs, hipStream_t s) {
#define SHG_KS(M, E, HB) launch_ks<M, E, A, HB, PL>(ka, gd, primes, lp, count, num_limbs, s)
  // The wave-exchange kernel at n = 2^12 (single LDS buffer: three 256-thread
  // workgroups per CU): 0.35 vs 0.52 us per ciphertext at L = 2, 1.28 vs 1.86
  // at L = 4.  At 2^13 and 2^14 one 512-thread workgroup per CU is resident
  // either way and keyswitch_kernel measured 5% and 6% faster (2^14, L = 8:
  // 10.55 vs 11.2 ms per 512 ciphertexts).
  if constexpr (PL && SHG_KS_N14 && std::is_same<typename A::V, double>::value) {
    if (logn == 14) return launch_ks_n14<A>(ka, gd, primes, lp, count, num_limbs, s);
  }
  if constexpr (std::is_same<A, F64KsKeyArith>::value) {
    // FP64-form keys at 2^12 (the wave-exchange kernel), 2^13 and 2^14.
    if (logn == 12) return launch_ks_wave<2, 0, A>(ka, gd, primes, lp, count, num_limbs, s);
    if (logn == 13) return SHG_KS(13, 4, 0);
    if (logn == 14) return SHG_KS(13, SHG_KS_LOGE, 1);
    return hipErrorInvalidValue;
  }
  if constexpr (PL && std::is_same<A, F64KsArith>::value && SHG_KS_WAVE) {
    if ((SHG_KS_WAVE_LOGN >> logn) & 1) {
      if (logn == 12) return launch_ks_wave<2, 0>(ka, gd, primes, lp, count, num_limbs, s);
      if (logn == 13) return launch_ks_wave<3, 0>(ka, gd, primes, lp, count, num_limbs, s);
      if constexpr (ks_planes_vc(14) == 32) {  // its plane layout at 2^14
        if (logn == 14) return launch_ks_wave<3, 1>(ka, gd, primes, lp, count, num_limbs, s);
      }
    }
  }
  if constexpr (!std::is_same<A, F64KsKeyArith>::value) {
  switch (logn) {
    case 4: return SHG_KS(4, 4, 0);
    case 5: return SHG_KS(5, 4, 0);
    case 6: return SHG_KS(6, 4, 0);
    case 7: return SHG_KS(7, 4, 0);
    case 8: return SHG_KS(8, 4, 0);
    case 9: return SHG_KS(9, 4, 0);
    case 10: return SHG_KS(10, 4, 0);
    case 11: return SHG_KS(11, 4, 0);
    case 12: return SHG_KS(12, 4, 0);
    case 13: return SHG_KS(13, 4, 0);
    case 14: return SHG_KS(13, SHG_KS_LOGE, 1);
    default: return hipErrorInvalidValue;
  }
  }
#undef SHG_KS
}

hipError_t launch_keyswitch(const KeySwitchArgs& ka, const GadgetDesc& gd, int variant,
                            const PrimeDev* primes, const LimbPrimes& lp, int count, int num_limbs,
                            int logn, hipStream_t s) {
  if (ka.key_f64 && !(variant == SHG_NTT_F64 && gd.small_digits && ka.planes && logn >= 12))
    return hipErrorInvalidValue;  // FP64-form keys are read by that kernel only
  if (variant == SHG_NTT_F64) {
    if (gd.small_digits) {
      if (ka.planes && ka.key_f64)
        return dispatch_ks<F64KsKeyArith, true>(logn, ka, gd, primes, lp, count, num_limbs, s);
      if (ka.planes)
        return dispatch_ks<F64KsArith, true>(logn, ka, gd, primes, lp, count, num_limbs, s);
      return dispatch_ks<F64KsArith, false>(logn, ka, gd, primes, lp, count, num_limbs, s);
    }
    return dispatch_ks<F64Arith, false>(logn, ka, gd, primes, lp, count, num_limbs, s);
  }
  return dispatch_ks<IntArith, false>(logn, ka, gd, primes, lp, count, num_limbs, s);
}

}  // namespace shg
