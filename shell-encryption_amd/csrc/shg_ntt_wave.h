// Forward NTT of a 2^(10+LW)-word slice with one workgroup-wide exchange per
// limb: 2^LW waves x 64 lanes x 16 words per lane.  Four register passes:
//
//   pass  stages (index bits)      registers   lanes            waves
//   1     top 4 bits               9+LW..6+LW  5..0             6+LW-1..6
//   2     6+LW-1..6 (LW stages)    9..6        5..0             top LW bits
//   3     5..2                     5..2        9..6, 1, 0       top LW bits
//   4     1..0                     9, 8, 1, 0  7..2             top LW bits
//
// Pass 1's words are loaded coalesced (lanes on the low bits) and its
// twiddles are uniform over the workgroup; the exchange to pass 2's layout
// goes through LDS between two barriers, after which each wave owns the
// words whose top LW bits are its index, so the exchanges to passes 3 and 4
// stay inside the wave (LDS slots the wave itself read, ordered by a wave
// barrier): between one limb's exchange and the next, waves run
// independently and their HBM phases overlap other waves' arithmetic.
// Pass 2's twiddles are uniform per wave.  Word i always sits at LDS slot
// lpad(i); lane, wave and register parts of an index occupy disjoint bits,
// so lpad(index) = lpad(lane/wave part) + lpad(register part).
#ifndef SHG_NTT_WAVE_H_
#define SHG_NTT_WAVE_H_

#include "shg_ntt_core.h"

namespace shg {

// Index bit of register bit k (0..3) and of lane bit k (0..5) in each layout.
template <int LW>
struct WaveLayout {
  static constexpr int LOGM = 10 + LW;
  __host__ __device__ static constexpr int reg(int pass, int k) {
    return pass == 1 ? LOGM - 4 + k : pass == 2 ? 6 + k : pass == 3 ? 2 + k
                                                                      : (k < 2 ? k : 6 + k);
  }
  __host__ __device__ static constexpr int lane(int pass, int k) {
    return pass <= 2 ? k : pass == 3 ? (k < 2 ? k : 4 + k) : 2 + k;
  }
  __host__ __device__ static constexpr int wave(int pass, int k) {
    return pass == 1 ? 6 + k : 10 + k;
  }
  // Register part of the index of register word e.
  __host__ __device__ static constexpr int reg_part(int pass, int e) {
    return ((e & 1) << reg(pass, 0)) | (((e >> 1) & 1) << reg(pass, 1)) |
           (((e >> 2) & 1) << reg(pass, 2)) | (((e >> 3) & 1) << reg(pass, 3));
  }
  // Lane and wave part of a thread's indices.
  __device__ static int thread_part(int pass, int lane_id, int wave_id) {
    int v = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) v |= ((lane_id >> k) & 1) << lane(pass, k);
#pragma unroll
    for (int k = 0; k < LW; ++k) v |= ((wave_id >> k) & 1) << wave(pass, k);
    return v;
  }
};

// Twiddles of one pass: stages on register bits kb = KHI down to KLO (index
// bits reg(P, kb)); global stage s = LOGM - 1 - index bit.  The pair (e,
// e | 1 << kb) uses psi_bitrev[2^s + (index >> (bit + 1))], which depends on
// the lane/wave part and on e's register bits above the stage's bit only:
// stage kb has 2^nhigh(kb) of them, stored from offset base(kb).
template <int LW, int P, int KHI, int KLO>
struct WavePassTw {
  using Lay = WaveLayout<LW>;
  __host__ __device__ static constexpr int nhigh(int kb) {
    int c = 0;
    for (int k = 0; k < 4; ++k) c += Lay::reg(P, k) > Lay::reg(P, kb) ? 1 : 0;
    return c;
  }
  // j-th register bit (increasing k) above stage kb's bit.
  __host__ __device__ static constexpr int high(int kb, int j) {
    for (int k = 0; k < 4; ++k)
      if (Lay::reg(P, k) > Lay::reg(P, kb) && j-- == 0) return k;
    return 0;
  }
  __host__ __device__ static constexpr int base(int kb) {
    int b = 0;
    for (int k = KHI; k > kb; --k) b += 1 << nhigh(k);
    return b;
  }
  static constexpr int N = base(KLO - 1);
  // Position of register word e's twiddle for stage kb.
  __host__ __device__ static constexpr int slot(int kb, int e) {
    int c = 0;
    for (int j = 0; j < nhigh(kb); ++j) c |= ((e >> high(kb, j)) & 1) << j;
    return base(kb) + c;
  }
  // Register part of the index for combination c of stage kb's high bits.
  __host__ __device__ static constexpr int combo_part(int kb, int c) {
    int v = 0;
    for (int j = 0; j < nhigh(kb); ++j) v |= ((c >> j) & 1) << Lay::reg(P, high(kb, j));
    return v;
  }
};

// The slice may be sub-block h of a limb with LOGM + HB index bits whose
// first HB stages are already applied: local stage s is global stage s + HB
// and its twiddle block carries h above the local part.
template <int LW, int P, int KHI, int KLO, int HB, class A>
__device__ __forceinline__ void wave_pass_tw_vec(double* tw, int tw_tp, int h, const A& ar) {
  using PT = WavePassTw<LW, P, KHI, KLO>;
  constexpr int LOGM = WaveLayout<LW>::LOGM;
#pragma unroll
  for (int kb = KHI; kb >= KLO; --kb) {
    const int bit = WaveLayout<LW>::reg(P, kb);
    const int s = LOGM - 1 - bit;
#pragma unroll
    for (int c = 0; c < (1 << PT::nhigh(kb)); ++c)
      tw[PT::base(kb) + c] = ar.fwd_tw_raw((unsigned)((1 << (s + HB)) + (h << s) +
                                                      ((PT::combo_part(kb, c) | tw_tp) >> (bit + 1))));
  }
}

#ifndef SHG_NTT_WAVE_SMEM
#define SHG_NTT_WAVE_SMEM 1
#endif

// tw_tp: the lane/wave part of the indices as far as it reaches above the
// pass's bits (0 for pass 1, the wave part for pass 2).  UNIFORM: tw_tp is
// the same across the wave, and the twiddles are read through the scalar
// cache (SMEM loads complete independently of vector loads in flight).
template <int LW, int P, int KHI, int KLO, bool UNIFORM = false, int HB = 0, class A>
__device__ __forceinline__ void wave_pass_tw(double* tw, int tw_tp, int h, const A& ar) {
  if constexpr (UNIFORM && SHG_NTT_WAVE_SMEM) {
    using PT = WavePassTw<LW, P, KHI, KLO>;
    constexpr int LOGM = WaveLayout<LW>::LOGM;
    using cptr = const double __attribute__((address_space(4)))*;
    const cptr ct = (cptr)(size_t)(const void*)ar.fw;
#pragma unroll
    for (int kb = KHI; kb >= KLO; --kb) {
      const int bit = WaveLayout<LW>::reg(P, kb);
      const int s = LOGM - 1 - bit;
#pragma unroll
      for (int c = 0; c < (1 << PT::nhigh(kb)); ++c)
        tw[PT::base(kb) + c] = ct[(unsigned)((1 << (s + HB)) + (h << s) +
                                             ((PT::combo_part(kb, c) | tw_tp) >> (bit + 1)))];
    }
    return;
  }
  wave_pass_tw_vec<LW, P, KHI, KLO, HB>(tw, tw_tp, h, ar);
}

template <int LW, int P, int KHI, int KLO, int HB = 0, class A>
__device__ __forceinline__ void wave_pass(double (&x)[16], const double* tw, const A& ar) {
  using PT = WavePassTw<LW, P, KHI, KLO>;
  constexpr int LOGM = WaveLayout<LW>::LOGM;
#pragma unroll
  for (int kb = KHI; kb >= KLO; --kb) {
    const int s = LOGM - 1 - WaveLayout<LW>::reg(P, kb) + HB;  // global stage
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      if (e & (1 << kb)) continue;
      ar.ct(x[e], x[e | (1 << kb)], ar.tw(tw[PT::slot(kb, e)]));
    }
    if (ar.fwd_reduce_at(s) && !(A::kSkipLastReduce && s == LOGM + HB - 1)) {
#pragma unroll
      for (int e = 0; e < 16; ++e) ar.fwd_reduce(x[e]);
    }
  }
}

// Inverse (Gentleman-Sande) passes run on the same layouts in the opposite
// order: pass 4's (index bits 0, 1), 3's (5..2), 2's (6+LW-1..6), 1's (the
// top 4 bits).  The stage on index bit b pairs word i (bit b clear) with
// i + 2^b and uses psi_inv_bitrev[n - n/2^b + (i >> (b + 1))], which depends
// on the same register bits above b as the forward twiddle of that bit, so
// WavePassTw's slots serve both.  UNIFORM as in wave_pass_tw.
//
// HB: the slice is sub-block h of a limb of 2^(LOGM + HB) words (its last HB
// stages run elsewhere); the twiddle index then carries h above the local
// part: n - n/2^b + (h << (LOGM - 1 - b)) + (local index >> (b + 1)).
template <int LW, int P, int KHI, int KLO, bool UNIFORM, int HB = 0, class A>
__device__ __forceinline__ void wave_pass_tw_inv(double* tw, int tw_tp, int h, const A& ar) {
  using PT = WavePassTw<LW, P, KHI, KLO>;
  constexpr int LOGM = WaveLayout<LW>::LOGM;
  constexpr int n = 1 << (LOGM + HB);
  if constexpr (UNIFORM && SHG_NTT_WAVE_SMEM) {
    using cptr = const double __attribute__((address_space(4)))*;
    const cptr ct = (cptr)(size_t)(const void*)ar.iw;
#pragma unroll
    for (int kb = KLO; kb <= KHI; ++kb) {
      const int bit = WaveLayout<LW>::reg(P, kb);
#pragma unroll
      for (int c = 0; c < (1 << PT::nhigh(kb)); ++c)
        tw[PT::base(kb) + c] = ct[(unsigned)(n - (n >> bit) + (h << (LOGM - 1 - bit)) +
                                             ((PT::combo_part(kb, c) | tw_tp) >> (bit + 1)))];
    }
  } else {
#pragma unroll
    for (int kb = KLO; kb <= KHI; ++kb) {
      const int bit = WaveLayout<LW>::reg(P, kb);
#pragma unroll
      for (int c = 0; c < (1 << PT::nhigh(kb)); ++c)
        tw[PT::base(kb) + c] =
            ar.inv_tw_raw((unsigned)(n - (n >> bit) + (h << (LOGM - 1 - bit)) +
                                     ((PT::combo_part(kb, c) | tw_tp) >> (bit + 1))));
    }
  }
}

// Stages on register bits KLO up to KRUN (default KHI; the twiddles are laid
// out for KHI..KLO); the reduction schedule follows the index bit (the
// inverse stage number).
template <int LW, int P, int KHI, int KLO, int KRUN = KHI, class A>
__device__ __forceinline__ void wave_pass_inv(double (&x)[16], const double* tw, const A& ar) {
  using PT = WavePassTw<LW, P, KHI, KLO>;
#pragma unroll
  for (int kb = KLO; kb <= KRUN; ++kb) {
    const int bit = WaveLayout<LW>::reg(P, kb);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      if (e & (1 << kb)) continue;
      ar.gs(x[e], x[e | (1 << kb)], ar.tw(tw[PT::slot(kb, e)]));
    }
    if (ar.inv_reduce_at(bit)) {
#pragma unroll
      for (int e = 0; e < 16; ++e) ar.inv_reduce(x[e]);
    }
  }
}

// The last inverse stage (the top index bit: register bit 3 of pass 1's
// layout, a single twiddle w) with an output scale s folded in:
// u' = (u + v) s, v' = (u - v)(w s), one product per word where the stage
// and a separate scaling take three per pair.
template <class A>
__device__ __forceinline__ void wave_last_inv_scaled(double (&x)[16], double w, double s,
                                                     const A& ar) {
  const double ws = f64_mulmod_h(w, s, ar.q, ar.qinv);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const double a = x[e] + x[e + 8];
    const double b = x[e] - x[e + 8];
    x[e] = f64_mulmod_h(a, s, ar.q, ar.qinv);
    x[e + 8] = f64_mulmod_h(b, ws, ar.q, ar.qinv);
  }
}

// Pass 1 of an inverse (top 4 bits) followed by the stores, with the
// scale (0: none, 1: n^-1, 2: n^-1 2^-64) folded into the last stage.
template <int LW, class A, class St>
__device__ __forceinline__ void wave_inv_last_pass_store(double (&x)[16], const double* tw4,
                                                         int scale, const A& ar, const St& st) {
  using T4 = WavePassTw<LW, 1, 3, 0>;
  if (scale == 0) {
    wave_pass_inv<LW, 1, 3, 0>(x, tw4, ar);
  } else {
    wave_pass_inv<LW, 1, 3, 0, 2>(x, tw4, ar);
    wave_last_inv_scaled(x, tw4[T4::slot(3, 0)], scale == 1 ? ar.ninv : ar.ninvx, ar);
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) st(e, ar.fwd_out(x[e]));
}

// Orders a wave's LDS writes before its following reads (other lanes').
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LW, int P>
__device__ __forceinline__ void wave_to_lds(const double (&x)[16], double* lds, int tp) {
  using Lay = WaveLayout<LW>;
  const int base = lpad(tp);
#pragma unroll
  for (int e = 0; e < 16; ++e) lds[base + cpad(Lay::reg_part(P, e))] = x[e];
}

template <int LW, int P>
__device__ __forceinline__ void wave_from_lds(double (&x)[16], const double* lds, int tp) {
  using Lay = WaveLayout<LW>;
  const int base = lpad(tp);
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = lds[base + cpad(Lay::reg_part(P, e))];
}

// Whole-workgroup transforms of one 2^14-word slice (LW = 4: 1024 threads)
// for kernels that run many differently sized jobs in one launch.  Words
// come from ld(p) and go to st(p, v) by slice position p: forward ld gives
// the policy's representation and st takes canonical words; inverse ld gives
// the raw residue and st takes inv_out(x, scale).  Both start with no LDS
// write before a barrier that other waves' earlier reads would need, and end
// with a barrier after their last LDS reads, so the caller may run any other
// LDS job right after (as after the register/LDS passes of shg_ntt_core.h).
// The forward slice's passes: x receives the transformed (unreduced) words,
// word e at slice position (return value) + reg_part(4, e).  No barrier
// after the last LDS reads.  PRELOADED: x already holds the input word of
// slice position thread_part(1, lane, wave) + reg_part(1, e) in x[e], and
// ld is not called.
template <int HB, bool PRELOADED = false, class A, class Ld>
__device__ __forceinline__ int wave_fwd_core(int h, const A& ar, double* lds, const Ld& ld,
                                             double (&x)[16]) {
  using Lay = WaveLayout<4>;
  using T1 = WavePassTw<4, 1, 3, 0>;
  using T2 = WavePassTw<4, 2, 3, 0>;
  using T3 = WavePassTw<4, 3, 3, 0>;
  using T4 = WavePassTw<4, 4, 1, 0>;
  int lane_o = threadIdx.x & 63;
  asm volatile("" : "+v"(lane_o));
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tp1 = Lay::thread_part(1, lane_o, wave);
  double tw1[T1::N];
  wave_pass_tw<4, 1, 3, 0, true, HB>(tw1, 0, h, ar);
  if constexpr (!PRELOADED) {
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = ld(tp1 + Lay::reg_part(1, e));
  }
  wave_pass<4, 1, 3, 0, HB>(x, tw1, ar);
  __builtin_amdgcn_sched_barrier(0);
  double tw2[T2::N];
  wave_pass_tw<4, 2, 3, 0, true, HB>(tw2, wave << 10, h, ar);
  __syncthreads();
  wave_to_lds<4, 1>(x, lds, tp1);
  __syncthreads();
  const int tp2 = Lay::thread_part(2, lane_o, wave);
  wave_from_lds<4, 2>(x, lds, tp2);
  wave_pass<4, 2, 3, 0, HB>(x, tw2, ar);
  const int tp3 = Lay::thread_part(3, lane_o, wave);
  __builtin_amdgcn_sched_barrier(0);
  double tw3[T3::N];
  wave_pass_tw<4, 3, 3, 0, false, HB>(tw3, tp3, h, ar);
  wave_to_lds<4, 2>(x, lds, tp2);
  wave_sync_lds();
  wave_from_lds<4, 3>(x, lds, tp3);
  wave_pass<4, 3, 3, 0, HB>(x, tw3, ar);
  const int tp4 = Lay::thread_part(4, lane_o, wave);
  __builtin_amdgcn_sched_barrier(0);
  double tw4[T4::N];
  wave_pass_tw<4, 4, 1, 0, false, HB>(tw4, tp4, h, ar);
  wave_to_lds<4, 3>(x, lds, tp3);
  wave_sync_lds();
  wave_from_lds<4, 4>(x, lds, tp4);
  pin_words(x);
  wave_pass<4, 4, 1, 0, HB>(x, tw4, ar);
  return tp4;
}

template <int HB, class A, class Ld, class St>
__device__ __forceinline__ void wave_fwd_slice(int h, const A& ar, double* lds, const Ld& ld,
                                               const St& st) {
  double x[16];
  const int tp4 = wave_fwd_core<HB>(h, ar, lds, ld, x);
#pragma unroll
  for (int e = 0; e < 16; ++e) st(tp4 + WaveLayout<4>::reg_part(4, e), ar.fwd_out(x[e]));
  __syncthreads();
}

// The inverse slice's passes up to pass 1's words in VGPRs: x holds word e
// of slice position (return value) + reg_part(1, e), tw4 pass 1's twiddles.
// The caller runs pass 1 (wave_pass_inv<4, 1, 3, 0> or
// wave_inv_last_pass_store) and a barrier before the next LDS write.
template <int HB, class A, class Ld>
__device__ __forceinline__ int wave_inv_core(int h, const A& ar, double* lds, const Ld& ld,
                                             double (&x)[16], double (&tw4)[15]) {
  using Lay = WaveLayout<4>;
  using T1 = WavePassTw<4, 4, 1, 0>;
  using T2 = WavePassTw<4, 3, 3, 0>;
  using T3 = WavePassTw<4, 2, 3, 0>;
  static_assert(WavePassTw<4, 1, 3, 0>::N == 15, "pass 1 has 15 twiddles");
  int lane_o = threadIdx.x & 63;
  asm volatile("" : "+v"(lane_o));
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tp4 = Lay::thread_part(4, lane_o, wave);
  double tw1[T1::N];
  wave_pass_tw_inv<4, 4, 1, 0, false, HB>(tw1, tp4, h, ar);
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = ar.load(ld(tp4 + Lay::reg_part(4, e)));
  wave_pass_inv<4, 4, 1, 0>(x, tw1, ar);
  __builtin_amdgcn_sched_barrier(0);
  const int tp3 = Lay::thread_part(3, lane_o, wave);
  double tw2[T2::N];
  wave_pass_tw_inv<4, 3, 3, 0, false, HB>(tw2, tp3, h, ar);
  wave_to_lds<4, 4>(x, lds, tp4);
  wave_sync_lds();
  wave_from_lds<4, 3>(x, lds, tp3);
  wave_pass_inv<4, 3, 3, 0>(x, tw2, ar);
  __builtin_amdgcn_sched_barrier(0);
  const int tp2 = Lay::thread_part(2, lane_o, wave);
  double tw3[T3::N];
  wave_pass_tw_inv<4, 2, 3, 0, true, HB>(tw3, wave << 10, h, ar);
  wave_to_lds<4, 3>(x, lds, tp3);
  wave_sync_lds();
  wave_from_lds<4, 2>(x, lds, tp2);
  wave_pass_inv<4, 2, 3, 0>(x, tw3, ar);
  __builtin_amdgcn_sched_barrier(0);
  const int tp1 = Lay::thread_part(1, lane_o, wave);
  wave_pass_tw_inv<4, 1, 3, 0, true, HB>(tw4, 0, h, ar);
  wave_to_lds<4, 2>(x, lds, tp2);
  __syncthreads();
  wave_from_lds<4, 1>(x, lds, tp1);
  return tp1;
}

template <int HB, class A, class Ld, class St>
__device__ __forceinline__ void wave_inv_slice(int h, const A& ar, double* lds, int scale,
                                               const Ld& ld, const St& st) {
  double x[16], tw4[15];
  const int tp1 = wave_inv_core<HB>(h, ar, lds, ld, x, tw4);
  wave_inv_last_pass_store<4>(x, tw4, scale, ar, [&](int e, uint64_t v) {
    st(tp1 + WaveLayout<4>::reg_part(1, e), v);
  });
  __syncthreads();
}

}  // namespace shg

#endif  // SHG_NTT_WAVE_H_
