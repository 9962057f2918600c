// Register/LDS building blocks of the negacyclic NTT on gfx950, shared by the
// transform kernels and the fused key-switch kernel.
//
// Semantics: ForwardNumberTheoreticTransform (Cooley-Tukey, natural order in,
// bit-reversed order out, twiddle psis_bitrev[2^s + block] at stage s) and
// InverseNumberTheoreticTransform (Gentleman-Sande, twiddle
// psis_inv_bitrev[n - n/2^i + block] at stage i, then times n^-1)
// (shell_encryption/dft_transformations.h:59-201).
//
// A workgroup owns a 2^LOGM-word slice in LDS; each thread keeps E = 2^LOGE
// words in VGPRs and runs up to LOGE radix-2 stages on them between LDS
// exchanges.  Two arithmetic policies, bit-exact with each other:
//  * IntArith: 64-bit Shoup products composed from 32-bit mul_lo/mul_hi;
//    values grow lazily by < 2q per stage (q < 2^50 leaves 14 spare bits),
//    one Barrett reduction at the end.
//  * F64Arith: exact products from FP64 FMA (q < 2^50), values kept balanced
//    and reduced every few stages.
#ifndef SHG_NTT_CORE_H_
#define SHG_NTT_CORE_H_

#include <hip/hip_runtime.h>

#include <type_traits>

#include "shg_device.h"

#ifndef SHG_STAGE_BARRIER
#define SHG_STAGE_BARRIER 0
#endif

namespace shg {


// Makes the words in x materialize in VGPRs at this point.  Placed after
// LDS reads whose slots other threads overwrite once a following barrier is
// passed: the volatile asm cannot move past that barrier and consumes the
// read values, so the reads cannot be sunk below it (the backend otherwise
// may sink loads whose only uses lie beyond the barrier).
template <class V, int E>
__device__ __forceinline__ void pin_words(V (&x)[E]) {
#pragma unroll
  for (int i = 0; i < E; ++i) asm volatile("" : "+v"(x[i]));
}

// One padding word per 16 keeps strided u64 LDS accesses spread over banks.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int cpad(int i) { return i + (i >> 4); }

// Padded LDS address of word (base + (e << PLO)) where base's bits below PLO
// come from the thread (lo) and e has R bits.  When the padding is linear
// over the two parts (PLO >= 4, or lo == 0 and base a multiple of 16) the
// address is one per-thread base plus a compile-time offset per e.
template <int PLO, int R>
struct LdsMap {
  static constexpr bool kLinear = (PLO >= 4) || (PLO == 0 && R >= 4);
  __device__ __forceinline__ static int base(int hi, int lo) {
    const int b = (hi << (PLO + R)) + lo;
    return kLinear ? lpad(b) : b;
  }
  __device__ __forceinline__ static int addr(int pbase, int e) {
    if constexpr (kLinear) return pbase + cpad(e << PLO);
    return lpad(pbase + (e << PLO));
  }
};

template <int LOGM>
constexpr int lds_words() {
  return (1 << LOGM) + ((1 << LOGM) >> 4);
}

// ---------------------------------------------------------------------------
// Arithmetic policies
// ---------------------------------------------------------------------------

struct IntArith {
  using V = uint64_t;
  // Whether the reduction scheduled after the last forward stage may be
  // left to the consumer (fwd_out or the key switch's prep reduce anyway).
  static constexpr bool kSkipLastReduce = false;
  uint64_t q, q2, bn;
  gptr<uint64_t> fw;
  gptr<uint64_t> iw;
  uint64_t ninv_c, ninv_cb, ninvx_c, ninvx_cb;
  // Lazy growth is < 2q per CT stage, so up to 15 stages stay below 2^64 for
  // q < 2^58; wider moduli (up to the reference's 2^62 limit) are reduced
  // after every stage.  GS outputs stay in [0, 2q).
  bool reduce_each;
  __device__ __forceinline__ explicit IntArith(const PrimeDev& P)
      : q(P.q), q2(2 * P.q), bn(P.bn), fw((gptr<uint64_t>)P.fw), iw((gptr<uint64_t>)P.iw),
        ninv_c(P.n_inv_c),
        ninv_cb(P.n_inv_cb), ninvx_c(P.n_inv_x_c), ninvx_cb(P.n_inv_x_cb),
        reduce_each(P.q >= (1ull << 58)) {}
  __device__ __forceinline__ bool fwd_reduce_at(int) const { return reduce_each; }
  __device__ __forceinline__ bool inv_reduce_at(int) const { return false; }
  struct Tw {
    uint64_t c, cb;
  };
  using TwRaw = Tw;  // (c, cb) as stored
  __device__ __forceinline__ TwRaw fwd_tw_raw(int i) const { return Tw{fw[2 * i], fw[2 * i + 1]}; }
  __device__ __forceinline__ TwRaw inv_tw_raw(int i) const { return Tw{iw[2 * i], iw[2 * i + 1]}; }
  __device__ __forceinline__ Tw tw(const TwRaw& r) const { return r; }
  __device__ __forceinline__ Tw fwd_tw(int i) const { return fwd_tw_raw(i); }
  __device__ __forceinline__ Tw inv_tw(int i) const { return inv_tw_raw(i); }
  __device__ __forceinline__ V load(uint64_t x) const { return x; }
  // Cooley-Tukey: values grow by less than 2q per stage.
  __device__ __forceinline__ void ct(V& u, V& v, const Tw& w) const {
    const uint64_t t = shoup_lazy(v, w.c, w.cb, q);
    v = u + q2 - t;
    u = u + t;
  }
  // One output of a Cooley-Tukey butterfly: u - w v when `low`, else u + w v.
  __device__ __forceinline__ V ct_half(V u, V v, const Tw& w, bool low) const {
    const uint64_t t = shoup_lazy(v, w.c, w.cb, q);
    return low ? u + q2 - t : u + t;
  }
  // Gentleman-Sande on [0, 2q) inputs, outputs in [0, 2q).
  __device__ __forceinline__ void gs(V& u, V& v, const Tw& w) const {
    const uint64_t s = u + v;
    const uint64_t d = u + q2 - v;
    u = csub(s, q2);
    v = shoup_lazy(d, w.c, w.cb, q);
  }
  __device__ __forceinline__ void fwd_reduce(V& x) const { x = reduce_full(x, q, bn); }
  __device__ __forceinline__ void inv_reduce(V&) const {}
  __device__ __forceinline__ uint64_t fwd_out(V x) const { return reduce_full(x, q, bn); }
  // mode 1: times n^-1 (InverseNumberTheoreticTransform); mode 2: times
  // n^-1 * 2^-64, i.e. the inverse transform followed by ExportInt.
  __device__ __forceinline__ uint64_t inv_out(V x, int mode) const {
    if (mode == 0) return fwd_out(x);
    return mode == 1 ? csub(shoup_lazy(x, ninv_c, ninv_cb, q), q)
                     : csub(shoup_lazy(x, ninvx_c, ninvx_cb, q), q);
  }

  // Key-switch accumulation (montgomery.h:446-465): digits enter the
  // transform as Montgomery representations and the key product is a REDC.
  using Acc = uint64_t;
  using Prep = uint64_t;
  // Transform outputs are reduced before the MAC; REDC accumulators stay
  // canonical.
  __host__ __device__ static constexpr bool ks_prep(int) { return true; }
  __host__ __device__ static constexpr int ks_acc_every(int) { return 1 << 30; }
  __device__ __forceinline__ V digit_in(uint64_t d, const PrimeDev& P) const {
    return mont_import(d, P);
  }
  __device__ __forceinline__ V digit_in32(uint32_t d, const PrimeDev& P) const {
    return mont_import(d, P);
  }
  // A digit of the limb of modulus qs that may exceed qs / 2: balanced lift.
  __device__ __forceinline__ V digit_in_bal(uint64_t d, uint64_t qs, const PrimeDev& P) const {
    return import_balanced(d, qs, P);
  }
  __device__ __forceinline__ Acc acc_in(uint64_t rep) const { return rep; }
  __device__ __forceinline__ Prep prep(V y) const { return fwd_out(y); }
  __device__ __forceinline__ void mac(Acc& a, Prep d, uint64_t key, const PrimeDev& P) const {
    a = mont_fma(a, d, key, P);
  }
  __device__ __forceinline__ void acc_reduce(Acc&) const {}
  __device__ __forceinline__ uint64_t acc_out(Acc a) const { return a; }
};

// FP64 twiddles read with buffer loads (F64Table).
#ifndef SHG_TW_BUFFER
#define SHG_TW_BUFFER 1
#endif

// Twiddle products take their quotient from fl(v*w) (f64_mulmod_h); the
// twiddle's w/q is then never formed.
#ifndef SHG_F64_HQ
#define SHG_F64_HQ 1
#endif
#if SHG_F64_HQ
#define SHG_F64_MULMOD(x, tw) f64_mulmod_h(x, (tw).w, q, qinv)
#else
#define SHG_F64_MULMOD(x, tw) f64_mulmod(x, (tw).w, (tw).wq, q)
#endif

struct F64Arith {
  using V = double;
  static constexpr bool kSkipLastReduce = true;
  double q, qinv, ninv, ninvq, ninvx, ninvxq;
  gptr<double> fw;
  gptr<double> iw;
#if SHG_TW_BUFFER
  F64Table fwt, iwt;
#endif
  __device__ __forceinline__ explicit F64Arith(const PrimeDev& P)
      : q(P.qd), qinv(P.qinv), ninv(P.n_inv_d), ninvq(P.n_inv_dq), ninvx(P.n_inv_x_d),
        ninvxq(P.n_inv_x_dq), fw((gptr<double>)P.fwd), iw((gptr<double>)P.iwd)
#if SHG_TW_BUFFER
        , fwt(P.fwd), iwt(P.iwd)
#endif
  {}
  struct Tw {
    double w, wq;
  };
  // Tables hold w only (8 bytes per twiddle).  With SHG_F64_HQ the products
  // take their quotient from fl(v*w) * fl(1/q) and wq is dead code; without
  // it wq = w * fl(1/q) is formed on use.  Either way three roundings, which
  // the bounds below absorb.
  using TwRaw = double;
#if SHG_TW_BUFFER
  __device__ __forceinline__ TwRaw fwd_tw_raw(int i) const { return fwt[(uint32_t)i]; }
  __device__ __forceinline__ TwRaw inv_tw_raw(int i) const { return iwt[(uint32_t)i]; }
#else
  __device__ __forceinline__ TwRaw fwd_tw_raw(int i) const { return fw[i]; }
  __device__ __forceinline__ TwRaw inv_tw_raw(int i) const { return iw[i]; }
#endif
  __device__ __forceinline__ Tw tw(TwRaw r) const { return Tw{r, r * qinv}; }
  __device__ __forceinline__ Tw fwd_tw(int i) const { return tw(fw[i]); }
  __device__ __forceinline__ Tw inv_tw(int i) const { return tw(iw[i]); }
  // Inputs are canonical (<= q < 2^50).
  __device__ __forceinline__ V load(uint64_t x) const { return u52_to_f64(x); }
  __device__ __forceinline__ void ct(V& u, V& v, const Tw& w) const {
    const double t = SHG_F64_MULMOD(v, w);
    v = u - t;
    u = u + t;
  }
  // The product a Cooley-Tukey butterfly forms before its sum and
  // difference: ct_side(u, v, w, s) == fma(s, ct_product(v, w), u).
  __device__ __forceinline__ V ct_product(V v, const Tw& w) const { return SHG_F64_MULMOD(v, w); }
  __device__ __forceinline__ V ct_half(V u, V v, const Tw& w, bool low) const {
    const double t = SHG_F64_MULMOD(v, w);
    return u + (low ? -t : t);
  }
  // One output of a Cooley-Tukey butterfly with the side chosen by a
  // uniform sign (+1: u + w v, -1: u - w v): one FMA, where a select of -t
  // or t costs a sign flip and a 64-bit select per word.
  __device__ __forceinline__ V ct_side(V u, V v, const Tw& w, double sgn) const {
    return fma(sgn, SHG_F64_MULMOD(v, w), u);
  }
  __device__ __forceinline__ void gs(V& u, V& v, const Tw& w) const {
    const double s = u + v;
    const double d = u - v;
    u = s;
    v = SHG_F64_MULMOD(d, w);
  }
  __device__ __forceinline__ void fwd_reduce(V& x) const { x = f64_reduce(x, q, qinv); }
  __device__ __forceinline__ void inv_reduce(V& x) const { x = f64_reduce(x, q, qinv); }
  __device__ __forceinline__ uint64_t fwd_out(V x) const {
    return f64_canon_u52(f64_reduce(x, q, qinv), q);
  }
  __device__ __forceinline__ uint64_t inv_out(V x, int mode) const {
    if (mode == 0) return fwd_out(x);
    const double y = mode == 1 ? f64_mulmod(x, ninv, ninvq, q) : f64_mulmod(x, ninvx, ninvxq, q);
    return f64_canon_u52(f64_reduce(y, q, qinv), q);
  }

  // Key-switch accumulation.  The digit value d (not its Montgomery form)
  // enters the transform, and the key's Montgomery representation k is used
  // as a plain residue: NTT(d) * k = NTT(d * 2^64) * k * 2^-64 mod q, which
  // is the reference's FusedMulAddInPlace of the imported digit.  Reduced
  // transform outputs (|d| <= q/2) give terms below 0.7 q, so the
  // accumulator is reduced every eight digits.
  using Acc = double;
  using Prep = double;
  __host__ __device__ static constexpr bool ks_prep(int) { return true; }
  __host__ __device__ static constexpr int ks_acc_every(int) { return 8; }
  // Digits are at most the exported residue, so below 2^50.
  __device__ __forceinline__ V digit_in(uint64_t d, const PrimeDev&) const {
    return u52_to_f64(d);
  }
  __device__ __forceinline__ V digit_in32(uint32_t d, const PrimeDev&) const { return (double)d; }
  // Balanced lift of a digit of the limb of modulus qs (< 2^50): d > qs / 2
  // enters as the exact negative d - qs, congruent to d modulo qs itself.
  __device__ __forceinline__ V digit_in_bal(uint64_t d, uint64_t qs, const PrimeDev&) const {
    return u52_to_f64(d) - (d > (qs >> 1) ? u52_to_f64(qs) : 0.0);
  }
  __device__ __forceinline__ Acc acc_in(uint64_t rep) const { return u52_to_f64(rep); }
  __device__ __forceinline__ Prep prep(V y) const { return f64_reduce(y, q, qinv); }
  __device__ __forceinline__ void mac(Acc& a, Prep d, uint64_t key, const PrimeDev&) const {
    const double w = u52_to_f64(key);
#if SHG_F64_HQ
    a += f64_mulmod_h(d, w, q, qinv);
#else
    a += f64_mulmod(d, w, w * qinv, q);
#endif
  }
  __device__ __forceinline__ void acc_reduce(Acc& a) const { a = f64_reduce(a, q, qinv); }
  __device__ __forceinline__ uint64_t acc_out(Acc a) const {
    return f64_canon_u52(f64_reduce(a, q, qinv), q);
  }
  // Reduction schedule (s, i: 0-based global stage after which to reduce).
  // For q < 2^50, f64_mulmod(x, w) with w < q returns |t| <= q/2 +
  // 1.5 * 2^-52 q |x| <= q/2 + 0.375 |x| (three roundings in
  // x * fl(w * fl(1/q))).  A CT stage maps a bound B to 1.375 B + q/2:
  // from canonical inputs (B = q) four stages give B < 7.1 q, and from a
  // reduction (B = q/2 + 2^-49 q) five stages give B < 7.7 q; both stay below
  // 8 q < 2^53, so every sum and product term is an exact integer.  A GS
  // stage at most doubles B: three stages from B = q, four from B = q/2.
  // A reduction due after the last forward stage is skipped
  // (kSkipLastReduce): the words then go to fwd_out or the key switch's
  // prep, which reduce any |x| < 2^53 themselves.
  __device__ static constexpr bool fwd_reduce_at(int s) {
    return s == 3 || (s > 3 && (s - 3) % 5 == 0);
  }
  __device__ static constexpr bool inv_reduce_at(int i) {
    return i == 2 || (i > 2 && (i - 2) % 4 == 0);
  }
};

// ---------------------------------------------------------------------------
// Register passes.  A pass covers R consecutive stages; a thread holds
// G = E / 2^R groups of 2^R words; group gamma = gg*T + t.
// ---------------------------------------------------------------------------

// Twiddles used by one pass, per group gg: stage j needs 2^j of them
// (blocks bh), stored at [gg][2^j - 1 + bh].
template <int LOGE, int R>
struct PassTw {
  static constexpr int G = (1 << LOGE) >> R;
  static constexpr int PER = (1 << R) - 1;
  static constexpr int N = G * PER;
};

// Forward pass: stages S0..S0+R-1 (global stage s toggles index bit
// LOGN-1-s).  The slice is sub-block `h` of a limb with LOGN bits; local
// indices have LOGM bits, so the twiddle block carries h in its top bits.
// Twiddles are loaded by fwd_pass_tw before the LDS exchange that precedes
// the pass, so that their latency overlaps it; in the first pass they are
// uniform across the workgroup and come through the scalar cache.
template <int LOGM, int LOGE, int S0, int R, int HB, class A>
__device__ __forceinline__ void fwd_pass_tw(typename A::TwRaw* tw, int t, int h, const A& ar) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int G = PassTw<LOGE, R>::G;
  constexpr int PER = PassTw<LOGE, R>::PER;
  constexpr int PLO = LOGM - S0 - R;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const int gam = gg * T + t;
    const int hi = (h << S0) + (gam >> PLO);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int s = S0 + j + HB;
#pragma unroll
      for (int bh = 0; bh < (1 << j); ++bh) {
        int idx = (1 << s) + (hi << j) + bh;
        if constexpr (S0 == 0 && G == 1) idx = __builtin_amdgcn_readfirstlane(idx);
        tw[gg * PER + (1 << j) - 1 + bh] = ar.fwd_tw_raw(idx);
      }
    }
  }
}

template <int LOGM, int LOGE, int S0, int R, int HB, class A>
__device__ __forceinline__ void fwd_pass(typename A::V (&x)[1 << LOGE],
                                         const typename A::TwRaw* tw, const A& ar) {
  constexpr int E = 1 << LOGE;
  constexpr int G = PassTw<LOGE, R>::G;
  constexpr int PER = PassTw<LOGE, R>::PER;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = S0 + j + HB;  // global stage
    const int eb = R - 1 - j;
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
#pragma unroll
      for (int bh = 0; bh < (1 << j); ++bh) {
        const auto w = ar.tw(tw[gg * PER + (1 << j) - 1 + bh]);
#pragma unroll
        for (int el = 0; el < (1 << eb); ++el) {
          const int e = (bh << (eb + 1)) | el;
          ar.ct(x[gg * (1 << R) + e], x[gg * (1 << R) + e + (1 << eb)], w);
        }
      }
    }
    if (ar.fwd_reduce_at(s) && !(A::kSkipLastReduce && s == LOGM + HB - 1)) {
#pragma unroll
      for (int i = 0; i < E; ++i) ar.fwd_reduce(x[i]);
    }
#if SHG_STAGE_BARRIER
    __builtin_amdgcn_sched_barrier(0);
#endif
  }
}

template <int LOGM, int LOGE, int S0, int R, class V>
__device__ __forceinline__ void fwd_to_lds(const V (&x)[1 << LOGE], V* lds, int t) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int G = (1 << LOGE) >> R;
  constexpr int PLO = LOGM - S0 - R;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const int gam = gg * T + t;
    const int pb = LdsMap<PLO, R>::base(gam >> PLO, gam & ((1 << PLO) - 1));
#pragma unroll
    for (int e = 0; e < (1 << R); ++e) lds[LdsMap<PLO, R>::addr(pb, e)] = x[gg * (1 << R) + e];
  }
}

template <int LOGM, int LOGE, int S0, int R, class V>
__device__ __forceinline__ void fwd_from_lds(V (&x)[1 << LOGE], const V* lds, int t) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int G = (1 << LOGE) >> R;
  constexpr int PLO = LOGM - S0 - R;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const int gam = gg * T + t;
    const int pb = LdsMap<PLO, R>::base(gam >> PLO, gam & ((1 << PLO) - 1));
#pragma unroll
    for (int e = 0; e < (1 << R); ++e) x[gg * (1 << R) + e] = lds[LdsMap<PLO, R>::addr(pb, e)];
  }
}

// Runs forward passes S0.. to LOGM-1 (recursive over passes).  With
// LAST_TO_LDS the final pass is written to LDS (natural layout, read back
// coalesced); otherwise it stays in VGPRs, where thread t holds, for each of
// its G groups gamma = gg*T + t, the 2^R consecutive words
// (gamma << R) .. (gamma << R) + 2^R - 1 (see FwdLast).
struct NoHook {
  template <class P>
  __device__ __forceinline__ void operator()(P) const {}
};

// Every exchange keeps word i at LDS slot lpad(i), and a pass reads and
// writes back the same words, so a thread's writes never land on a slot
// that another thread still has to read in that pass: one barrier (after
// the writes) separates consecutive passes.  Only a last pass that leaves
// its words in VGPRs waits after its reads, so that the caller may write
// LDS again.  The twiddles of pass p + 1 are issued before pass p's
// exchange, so their latency overlaps its writes and barrier.
// With DB consecutive exchanges alternate between lds0 and lds1 and no last
// pass waits after its reads (the caller must continue the alternation
// across calls: swap the buffers after a call with an odd number of
// exchanges).  hook(integral_constant<S0>) runs right after the twiddle
// loads of the pass starting at stage S0 are issued, for callers that issue
// loads of their own there (vector loads complete in issue order, so a load
// issued after the twiddles does not delay them).
template <int LOGM, int LOGE, int S0, class A, bool LAST_TO_LDS = true, int HB = 0,
          bool DB = false>
struct FwdPasses {
  static constexpr int R = (LOGM - S0) < LOGE ? (LOGM - S0) : LOGE;
  static constexpr int P = S0 / LOGE;  // pass index (all passes but the last have R = LOGE)
  static constexpr bool kLast = S0 + R >= LOGM;
  using Tw = typename A::TwRaw[PassTw<LOGE, R>::N];
  template <class Hook = NoHook>
  __device__ __forceinline__ static void run(typename A::V (&x)[1 << LOGE], typename A::V* lds0,
                                             typename A::V* lds1, int t, int h, const A& ar,
                                             const Hook& hook = Hook()) {
    Tw tw;
    fwd_pass_tw<LOGM, LOGE, S0, R, HB>(tw, t, h, ar);
    hook(std::integral_constant<int, S0>());
    run_tw(x, tw, lds0, lds1, t, h, ar, hook);
  }
  // The same with this pass's twiddles already loaded.
  template <class Hook>
  __device__ __forceinline__ static void run_tw(typename A::V (&x)[1 << LOGE], const Tw& tw,
                                                typename A::V* lds0, typename A::V* lds1, int t,
                                                int h, const A& ar, const Hook& hook) {
    if constexpr (S0 > 0) {
      fwd_from_lds<LOGM, LOGE, S0, R>(x, (DB && ((P - 1) & 1)) ? lds1 : lds0, t);
      if constexpr (kLast && !LAST_TO_LDS) {
        pin_words(x);
        if constexpr (!DB) __syncthreads();
      }
    }
    fwd_pass<LOGM, LOGE, S0, R, HB>(x, tw, ar);
    if constexpr (!kLast) {
      using Next = FwdPasses<LOGM, LOGE, S0 + R, A, LAST_TO_LDS, HB, DB>;
      typename Next::Tw tw2;
      fwd_pass_tw<LOGM, LOGE, S0 + R, Next::R, HB>(tw2, t, h, ar);
      hook(std::integral_constant<int, S0 + R>());
      fwd_to_lds<LOGM, LOGE, S0, R>(x, (DB && (P & 1)) ? lds1 : lds0, t);
      __syncthreads();
      Next::run_tw(x, tw2, lds0, lds1, t, h, ar, hook);
    } else if constexpr (LAST_TO_LDS) {
      fwd_to_lds<LOGM, LOGE, S0, R>(x, (DB && (P & 1)) ? lds1 : lds0, t);
      __syncthreads();
    }
  }
  __device__ __forceinline__ static void run(typename A::V (&x)[1 << LOGE], typename A::V* lds,
                                             int t, int h, const A& ar) {
    run(x, lds, lds, t, h, ar);
  }
  // LDS exchanges made by one run.
  static constexpr int kPasses = (LOGM - S0 + LOGE - 1) / LOGE;
  static constexpr int kExchanges = kPasses - 1 + (LAST_TO_LDS ? 1 : 0);
};

// Shape of the last forward pass: R_LAST stages, G_LAST groups per thread.
template <int LOGM, int LOGE, int S0 = 0>
struct FwdLast {
  static constexpr int R = (LOGM - S0) < LOGE ? (LOGM - S0) : LOGE;
  static constexpr int R_LAST = (S0 + R >= LOGM) ? R : FwdLast<LOGM, LOGE, S0 + R>::R_LAST;
  static constexpr int G_LAST = (1 << LOGE) >> R_LAST;
};
template <int LOGM, int LOGE>
struct FwdLast<LOGM, LOGE, LOGM> {
  static constexpr int R = 0;
  static constexpr int R_LAST = 0;
  static constexpr int G_LAST = 0;
};

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// Two 64-bit words from one 16-byte buffer load, assembled from the load's
// four dwords into plain scalars.  Passing them through a two-lane 64-bit
// vector and converting a lane to double has let the optimizer (hipcc of
// ROCm 7.2) shrink the load to its first 8 bytes, leaving the second word
// unread; scalars keep the 16-byte load and cost no instruction.
struct U64Pair {
  uint64_t x, y;
};
__device__ __forceinline__ U64Pair buffer_load_u64x2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  U64Pair out;
  out.x = (uint64_t)v.x | ((uint64_t)v.y << 32);
  out.y = (uint64_t)v.z | ((uint64_t)v.w << 32);
  return out;
}

// Inverse pass: stages I0..I0+R-1 (stage i toggles local bit i).  Local
// stage i of sub-block h is global stage i with block index carrying h.
// Stage j of the pass needs 2^(R-1-j) twiddles per group, stored at
// [gg][2^R - 2^(R-j) + bh].
template <int LOGM, int LOGE, int I0, int R, int LOGN, class A>
__device__ __forceinline__ void inv_pass_tw(typename A::TwRaw* tw, int t, int h, const A& ar) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int G = PassTw<LOGE, R>::G;
  constexpr int PER = PassTw<LOGE, R>::PER;
  constexpr int n = 1 << LOGN;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const int gam = gg * T + t;
    const int hi = (h << (LOGM - I0 - R)) + (gam >> I0);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = I0 + j;
#pragma unroll
      for (int bh = 0; bh < (1 << (R - 1 - j)); ++bh)
        tw[gg * PER + (1 << R) - (1 << (R - j)) + bh] =
            ar.inv_tw_raw(n - (n >> i) + (hi << (R - 1 - j)) + bh);
    }
  }
}

template <int LOGM, int LOGE, int I0, int R, class A>
__device__ __forceinline__ void inv_pass(typename A::V (&x)[1 << LOGE],
                                         const typename A::TwRaw* tw,
                                         const A& ar) {
  constexpr int E = 1 << LOGE;
  constexpr int G = PassTw<LOGE, R>::G;
  constexpr int PER = PassTw<LOGE, R>::PER;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = I0 + j;
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
#pragma unroll
      for (int bh = 0; bh < (1 << (R - 1 - j)); ++bh) {
        const auto w = ar.tw(tw[gg * PER + (1 << R) - (1 << (R - j)) + bh]);
#pragma unroll
        for (int el = 0; el < (1 << j); ++el) {
          const int e = (bh << (j + 1)) | el;
          ar.gs(x[gg * (1 << R) + e], x[gg * (1 << R) + e + (1 << j)], w);
        }
      }
    }
    if (ar.inv_reduce_at(i)) {
#pragma unroll
      for (int k = 0; k < E; ++k) ar.inv_reduce(x[k]);
    }
#if SHG_STAGE_BARRIER
    __builtin_amdgcn_sched_barrier(0);
#endif
  }
}

template <int LOGM, int LOGE, int I0, int R, class V>
__device__ __forceinline__ void inv_to_lds(const V (&x)[1 << LOGE], V* lds, int t) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int G = (1 << LOGE) >> R;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const int gam = gg * T + t;
    const int pb = LdsMap<I0, R>::base(gam >> I0, gam & ((1 << I0) - 1));
#pragma unroll
    for (int e = 0; e < (1 << R); ++e) lds[LdsMap<I0, R>::addr(pb, e)] = x[gg * (1 << R) + e];
  }
}

template <int LOGM, int LOGE, int I0, int R, class V>
__device__ __forceinline__ void inv_from_lds(V (&x)[1 << LOGE], const V* lds, int t) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int G = (1 << LOGE) >> R;
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    const int gam = gg * T + t;
    const int pb = LdsMap<I0, R>::base(gam >> I0, gam & ((1 << I0) - 1));
#pragma unroll
    for (int e = 0; e < (1 << R); ++e) x[gg * (1 << R) + e] = lds[LdsMap<I0, R>::addr(pb, e)];
  }
}

// Inverse passes from I0; the last pass is left in registers.
// With FIRST_IN_REGS the caller has already loaded the first pass's words
// into VGPRs: thread t holds, for each group gamma = gg*T + t, the 2^R
// consecutive words (gamma << R) + e (the first pass toggles the low bits).
// As in FwdPasses, words stay at slot lpad(i), so only the last pass waits
// after its reads (the caller may then write LDS at once), and the next
// pass's twiddles are issued before an exchange.  hook(integral_constant<I0>)
// runs right after the twiddle loads of the pass starting at I0.
template <int LOGM, int LOGE, int I0, class A, bool FIRST_IN_REGS = false, int LOGN = LOGM>
struct InvPasses {
  static constexpr int R = (LOGM - I0) < LOGE ? (LOGM - I0) : LOGE;
  using Tw = typename A::TwRaw[PassTw<LOGE, R>::N];
  template <class Hook = NoHook>
  __device__ __forceinline__ static void run(typename A::V (&x)[1 << LOGE], typename A::V* lds,
                                             int t, int h, const A& ar,
                                             const Hook& hook = Hook()) {
    Tw tw;
    inv_pass_tw<LOGM, LOGE, I0, R, LOGN>(tw, t, h, ar);
    hook(std::integral_constant<int, I0>());
    run_tw(x, tw, lds, t, h, ar, hook);
  }
  template <class Hook>
  __device__ __forceinline__ static void run_tw(typename A::V (&x)[1 << LOGE], const Tw& tw,
                                                typename A::V* lds, int t, int h, const A& ar,
                                                const Hook& hook) {
    if constexpr (!(FIRST_IN_REGS && I0 == 0)) {
      inv_from_lds<LOGM, LOGE, I0, R>(x, lds, t);
      if constexpr (I0 + R >= LOGM) {
        pin_words(x);
        __syncthreads();
      }
    }
    inv_pass<LOGM, LOGE, I0, R>(x, tw, ar);
    if constexpr (I0 + R < LOGM) {
      using Next = InvPasses<LOGM, LOGE, I0 + R, A, false, LOGN>;
      typename Next::Tw tw2;
      inv_pass_tw<LOGM, LOGE, I0 + R, Next::R, LOGN>(tw2, t, h, ar);
      hook(std::integral_constant<int, I0 + R>());
      inv_to_lds<LOGM, LOGE, I0, R>(x, lds, t);
      __syncthreads();
      Next::run_tw(x, tw2, lds, t, h, ar, hook);
    }
  }
};

template <int LOGM, int LOGE, int I0, class A>
struct InvLast {
  static constexpr int R = (LOGM - I0) < LOGE ? (LOGM - I0) : LOGE;
  static constexpr int value = (I0 + R >= LOGM) ? I0 : InvLast<LOGM, LOGE, I0 + R, A>::value;
};
template <int LOGM, int LOGE, class A>
struct InvLast<LOGM, LOGE, LOGM, A> {
  static constexpr int value = LOGM;
};

}  // namespace shg

#endif  // SHG_NTT_CORE_H_
