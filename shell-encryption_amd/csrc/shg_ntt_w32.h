// Transforms of 2^14-word limbs with 32 words per lane, sized so that two
// workgroups share a CU: 512 threads (8 waves x 64 lanes), at most 128
// VGPRs, and 76 KiB of LDS (68 KiB of exchange slots, 8 KiB of twiddles).  While one workgroup loads or stores its limb
// the other computes, which one workgroup per CU holding a whole limb in
// 136 KiB of LDS (the wave-exchange kernels) cannot do.
//
// Three register passes of five index bits, two LDS exchanges:
//
//   layout  register bits 0..4   lane bits 0..5        wave bits 0..2
//   1       index bits 9..13     0..5                  6..8
//   2       4..8                 0..3, 9, 10           11..13
//   3       0..4                 5..10                 11..13
//
// Forward (Cooley-Tukey, stage s on index bit 13 - s): layout 1 is loaded
// coalesced and runs stages 0-4, whose twiddles are uniform over the
// workgroup (read through the scalar cache); exchange 1 goes to layout 2
// through LDS in two rounds over half the limb each: round r moves the
// words with index bit 13 = r, which layout 2's waves 4r..4r+3 own, so a
// wave of the second half still holds its round-1 words while it reads
// round 0 (48 words live at most); stages 5-9 follow with per-lane twiddles
// (contiguous runs of 1..16 words); exchange 2 stays inside each wave and
// moves in round r the words with index bit 4 = r, a register bit of both
// layouts; stages 10-13 run on layout 3, whose lane owns 32 consecutive
// words, stored with 16-byte writes.  The inverse (Gentleman-Sande, stage on
// index bit b for b = 0..13) runs the same layouts in the opposite order.
//
// The last forward pass leaves each lane 32 consecutive words; a third
// exchange inside each wave (rounds on index bit 0) turns them into the
// store layout S (register bits: index bits 0, 7..10; lanes: 1..6), in
// which every 16-byte store instruction of a wave covers 1 KiB of
// consecutive words.  The inverse loads layout S the same way.
//
// Twiddles: stages 0-4 read through the scalar cache; the forward's
// stages 5-9 use psi_bitrev[32..1024) and the inverse's bits 4-8 use
// psi_inv_bitrev[n - n/16 .. n - 32), 992 words either way, copied into LDS
// once per workgroup.  The per-lane runs of the last pass come from loads
// issued at the start of the kernel: whole runs for the short ones, and
// for the long ones the run's first twiddle times table entries that are
// uniform over the workgroup: psi_bitrev[2^s + 2^k h + c] =
// psi_bitrev[2^s + 2^k h] psi_bitrev[c] for 0 < c < 2^k (both are psi raised
// to the bit-reversed index, and the bits of 2^k h and c do not overlap),
// and in the inverse table the factor for c in [2^m, 2^(m+1)) is
// psi_inv_bitrev[n - 3 2^m + c].
//
// The slot of word i in exchange 1 is cpad(i mod 2^13); in exchanges 2
// and 3 it lies inside wave w's region of 1088 words: exchange 2 at
// cpad(i's bits 0-3 and 5-10), exchange 3 at cpad(i's bits 1-10).  Each
// address is a per-thread base plus a compile-time offset per register.
#ifndef SHG_NTT_W32_H_
#define SHG_NTT_W32_H_

#include "shg_ntt_core.h"
#include "shg_ntt_wave.h"

namespace shg {
namespace w32 {

constexpr int kLogN = 14;
constexpr int kThreads = 512;
constexpr int kXWords = 8704;        // exchange region: cpad(2^13) = 8 x cpad(2^10)
constexpr int kTwWords = 992;        // the middle pass's twiddles
constexpr int kLdsWords = kXWords + kTwWords;

__host__ __device__ constexpr int idx1(int w, int l, int e) { return (e << 9) | (w << 6) | l; }
__host__ __device__ constexpr int idx2(int w, int l, int e) {
  return (e << 4) | (l & 15) | ((l >> 4) << 9) | (w << 11);
}
__host__ __device__ constexpr int idx3(int w, int l, int e) { return e | (l << 5) | (w << 11); }

// Per-thread bases of the exchanges' slots:
//   exchange 1, layout 1: cpad(64 w + l) + 544 (e mod 16)
//   exchange 1, layout 2: cpad(layout-2 index without e and bit 13) + 17 e
//   exchange 2, layout 2: 1088 w + cpad((l mod 16) + 256 (l >> 4)) + 17 (e >> 1)
//   exchange 2, layout 3: 1088 w + 17 l + (e mod 16)
//   exchange 3, layout 3: 1088 w + 17 l + (e >> 1)
//   exchange 3, layout S: 1088 w + cpad(l) + 68 k (k: index bits 7..10)
struct Slots {
  int x1_l1, x1_l2, x2_l2, x2_l3, x3_s;
  __device__ __forceinline__ Slots(int w, int l) {
    x1_l1 = cpad((w << 6) | l);
    x1_l2 = cpad((l & 15) | ((l >> 4) << 9) | ((w & 3) << 11));
    x2_l2 = w * 1088 + cpad((l & 15) | ((l >> 4) << 8));
    x2_l3 = w * 1088 + 17 * l;
    x3_s = w * 1088 + cpad(l);
  }
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff,
                                           0x00020000);
}
__device__ __forceinline__ uint64_t ld64(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ U64Pair ld128(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return buffer_load_u64x2(r, voff, soff);
}
__device__ __forceinline__ void st64(__amdgpu_buffer_rsrc_t r, uint64_t v, int voff, int soff) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void st128(__amdgpu_buffer_rsrc_t r, U64Pair v, int voff, int soff) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  u32x4 d;
  d.x = (unsigned)v.x;
  d.y = (unsigned)(v.x >> 32);
  d.z = (unsigned)v.y;
  d.w = (unsigned)(v.y >> 32);
  __builtin_amdgcn_raw_buffer_store_b128(d, r, voff, soff, 0);
}

// CNT consecutive twiddles from index `start` (even when CNT > 1) of an FP64
// table read through buffer descriptor r.
template <int CNT>
__device__ __forceinline__ void tw_run(__amdgpu_buffer_rsrc_t r, int start, double* tw) {
  if constexpr (CNT == 1) {
    tw[0] = __builtin_bit_cast(double, ld64(r, start * 8, 0));
  } else {
#pragma unroll
    for (int k = 0; k < CNT / 2; ++k) {
      const U64Pair v = ld128(r, start * 8 + 16 * k, 0);
      tw[2 * k] = __builtin_bit_cast(double, v.x);
      tw[2 * k + 1] = __builtin_bit_cast(double, v.y);
    }
  }
}

// A run of CNT twiddles from its first one: out[c] = b * v[c] for c > 0.
template <int CNT, class A>
__device__ __forceinline__ void tw_from_base(double b, const double* v, double* out, const A& ar) {
  out[0] = b;
#pragma unroll
  for (int c = 1; c < CNT; ++c) out[c] = f64_mulmod_h(b, v[c], ar.q, ar.qinv);
}

// Forward stage on register bit KB of a layout whose register bits sit on
// consecutive index bits: the pair (e, e | 2^KB) takes tw[e >> (KB + 1)].
template <int KB, class A>
__device__ __forceinline__ void ct_stage(double (&x)[32], const double* tw, const A& ar) {
#pragma unroll
  for (int e = 0; e < 32; ++e) {
    if (e & (1 << KB)) continue;
    ar.ct(x[e], x[e | (1 << KB)], ar.tw(tw[e >> (KB + 1)]));
  }
}
template <int KB, class A>
__device__ __forceinline__ void gs_stage(double (&x)[32], const double* tw, const A& ar) {
#pragma unroll
  for (int e = 0; e < 32; ++e) {
    if (e & (1 << KB)) continue;
    ar.gs(x[e], x[e | (1 << KB)], ar.tw(tw[e >> (KB + 1)]));
  }
}

template <class A, int S>
__device__ __forceinline__ void fwd_reduce_after(double (&x)[32], const A& ar) {
  if constexpr (A::fwd_reduce_at(S) && !(A::kSkipLastReduce && S == kLogN - 1)) {
#pragma unroll
    for (int e = 0; e < 32; ++e) ar.fwd_reduce(x[e]);
  }
}
template <class A, int B>
__device__ __forceinline__ void inv_reduce_after(double (&x)[32], const A& ar) {
  if constexpr (A::inv_reduce_at(B)) {
#pragma unroll
    for (int e = 0; e < 32; ++e) ar.inv_reduce(x[e]);
  }
}

// Forward transform of one limb by one workgroup, FP64 arithmetic.  The
// caller provides the LDS (kLdsWords words) and the input/output policy:
// io.load(v1, e) returns word 512 e + 64 wave + lane (v1 = that word's
// byte offset for e = 0) as a double of magnitude at most q, and
// io.store(vs, k, lo, hi, aux) receives the unreduced outputs (|x| < 8q) of
// index 2 lane + 128 k + 2048 wave and the one after it (vs = the byte
// offset of the first for k = 0), and with IO::kAux the value io.aux(vs, k)
// returned, loaded ahead of the store (IO::Aux is its type).
template <class A, class IO>
__device__ __forceinline__ void fwd_limb(const A& ar, const PrimeDev& P, double* lds, const IO& io) {
  using cptr = const double __attribute__((address_space(4)))*;
  double* t2 = lds + kXWords;  // psi_bitrev[32 + j] at t2[j]
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const __amdgpu_buffer_rsrc_t twr = rsrc(P.fwd);
  const cptr ct = (cptr)(size_t)(const void*)P.fwd;
  const Slots sl(wave, lane);
  const int hi3 = lane | (wave << 6);  // index bits 5..13 of layout 3

  // The middle pass's twiddle block, loaded first (loads complete in
  // order) and copied into LDS below.
  const double t2a = __builtin_bit_cast(double, ld64(twr, (32 + t) * 8, 0));
  const double t2b = t < kTwWords - 512 ? __builtin_bit_cast(double, ld64(twr, (544 + t) * 8, 0)) : 0.0;
  // Layout 1: word e of the thread is index 512 e + 64 wave + lane.  The
  // words are loaded in stage-0 pairs (e, e + 16), so that the first
  // butterflies wait for the first loads only.
  double x[32];
  {
    const int v1 = ((wave << 6) | lane) * 8;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      x[e] = io.load(v1, e);
      x[e + 16] = io.load(v1, e + 16);
    }
  }
  t2[t] = t2a;
  if (t < kTwWords - 512) t2[512 + t] = t2b;
  // Stages 0-4 on register bits 4..0: twiddle psi_bitrev[2^s + (e >> (kb + 1))],
  // uniform over the workgroup.
  {
    auto stage = [&](auto sc) {
      constexpr int s = decltype(sc)::value;
      double tw[1 << s];
#pragma unroll
      for (int c = 0; c < (1 << s); ++c) tw[c] = ct[(1 << s) + c];
      ct_stage<4 - s>(x, tw, ar);
      fwd_reduce_after<A, s>(x, ar);
    };
    stage(std::integral_constant<int, 0>{});
    stage(std::integral_constant<int, 1>{});
    stage(std::integral_constant<int, 2>{});
    stage(std::integral_constant<int, 3>{});
    stage(std::integral_constant<int, 4>{});
  }
  // Exchange 1, round r: the words with index bit 13 = r (registers 16 r ..
  // 16 r + 15), read by the waves 4 r .. 4 r + 3 that own them in layout 2.
  // Its first barrier also publishes the twiddle block.
  double y[32];
#pragma unroll
  for (int e = 0; e < 16; ++e) lds[sl.x1_l1 + 544 * e] = x[e];
  __syncthreads();
  if (wave < 4) {
#pragma unroll
    for (int e = 0; e < 32; ++e) y[e] = lds[sl.x1_l2 + 17 * e];
    pin_words(y);
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 16; ++e) lds[sl.x1_l1 + 544 * e] = x[16 + e];
  __syncthreads();
  if (wave >= 4) {
#pragma unroll
    for (int e = 0; e < 32; ++e) y[e] = lds[sl.x1_l2 + 17 * e];
    pin_words(y);
  }
  // The last pass's twiddle runs, wanted after the middle pass: stages 10
  // and 11 whole, the first twiddle of stage 13.
  double tw10[2], tw11[4];
  tw_run<2>(twr, 1024 + (hi3 << 1), tw10);
  tw_run<4>(twr, 2048 + (hi3 << 2), tw11);
  const double b13 = __builtin_bit_cast(double, ld64(twr, (8192 + (hi3 << 4)) * 8, 0));
  __builtin_amdgcn_sched_barrier(0);
  // Stages 5-9 on register bits 4..0 of layout 2: twiddle psi_bitrev[2^s +
  // (hi << (4 - kb)) + (e >> (kb + 1))] with hi = index bits 9..13, from LDS.
  {
    const int hi = (lane >> 4) | (wave << 2);
    auto stage = [&](auto sc) {
      constexpr int s = decltype(sc)::value;
      constexpr int kb = 9 - s;
      double tw[1 << (4 - kb)];
      const double* src = t2 + ((1 << s) - 32) + (hi << (4 - kb));
#pragma unroll
      for (int c = 0; c < (1 << (4 - kb)); ++c) tw[c] = src[c];
      ct_stage<kb>(y, tw, ar);
      fwd_reduce_after<A, s>(y, ar);
    };
    stage(std::integral_constant<int, 5>{});
    stage(std::integral_constant<int, 6>{});
    stage(std::integral_constant<int, 7>{});
    stage(std::integral_constant<int, 8>{});
    stage(std::integral_constant<int, 9>{});
  }
  // Exchange 2 writes into every wave's region, which waves 4-7 read in
  // round 1 of exchange 1 (the middle pass runs before this barrier, so
  // waves 0-3 do not wait for those reads).
  __syncthreads();
  // Exchange 2 (inside the wave), round r: the words with index bit 4 = r
  // (layout 2: even or odd registers; layout 3: registers 16 r .. 16 r + 15).
  double z[32];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int e = r; e < 32; e += 2) lds[sl.x2_l2 + 17 * (e >> 1)] = y[e];
    wave_sync_lds();
#pragma unroll
    for (int e = 0; e < 16; ++e) z[16 * r + e] = lds[sl.x2_l3 + e];
    pin_words(z);
    wave_sync_lds();
  }
  // Stages 10-13 on register bits 3..0 of layout 3: twiddle psi_bitrev[2^s +
  // (hi3 << (4 - kb)) + (e >> (kb + 1))].  Stage 12's run is loaded here,
  // two stages ahead of its use (earlier it would not fit in the VGPRs).
  double tw12[8];
  tw_run<8>(twr, 4096 + (hi3 << 3), tw12);
  ct_stage<3>(z, tw10, ar);
  fwd_reduce_after<A, 10>(z, ar);
  ct_stage<2>(z, tw11, ar);
  fwd_reduce_after<A, 11>(z, ar);
  {
    double v[16];
#pragma unroll
    for (int c = 1; c < 16; ++c) v[c] = ct[c];
    double tw[16];
    ct_stage<1>(z, tw12, ar);
    fwd_reduce_after<A, 12>(z, ar);
    tw_from_base<16>(b13, v, tw, ar);
    ct_stage<0>(z, tw, ar);
    fwd_reduce_after<A, 13>(z, ar);
  }
  // A store policy that reads the output positions first (IO::kAux) gets
  // its words for k < 8 loaded now and those for k + 8 while k is stored.
  const int vs = ((lane << 1) | (wave << 11)) * 8;
  typename IO::Aux aux[16];
  if constexpr (IO::kAux) {
#pragma unroll
    for (int k = 0; k < 8; ++k) aux[k] = io.aux(vs, k);
  }
  // Exchange 3 (inside the wave), round r: the words with index bit 0 = r,
  // to layout S: lane l, register k holds index r + 2 l + 128 k + 2048 wave.
  double so[2][16];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int e = r; e < 32; e += 2) lds[sl.x2_l3 + (e >> 1)] = z[e];
    wave_sync_lds();
#pragma unroll
    for (int k = 0; k < 16; ++k) so[r][k] = lds[sl.x3_s + 68 * k];
    if (r == 0) {
      pin_words(so[0]);
      wave_sync_lds();
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if constexpr (IO::kAux) {
      if (k < 8) aux[k + 8] = io.aux(vs, k + 8);
    }
    io.store(vs, k, so[0][k], so[1][k], aux[k]);
  }
}

// Inverse transform of one limb by one workgroup, the output times sc
// (folded into the last stage; sc = 1 reduces exactly).  io.load(vs, k,
// lo, hi) gives the input words of index 2 lane + 128 k + 2048 wave and the
// one after it as doubles of magnitude at most q; io.store(v1, e, x)
// receives the output of index 512 e + 64 wave + lane (|x| < 2q), or with
// IO::kAll io.store_all(v1, x) receives all 32 at once.
template <class A, class IO>
__device__ __forceinline__ void inv_limb(const A& ar, const PrimeDev& P, double* lds, double sc,
                                         const IO& io) {
  using cptr = const double __attribute__((address_space(4)))*;
  constexpr int n = 1 << kLogN;
  constexpr int kT2Base = n - n / 16;  // first twiddle of bit 4
  double* t2 = lds + kXWords;  // psi_inv_bitrev[kT2Base + j] at t2[j]
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const __amdgpu_buffer_rsrc_t twr = rsrc(P.iwd);
  const cptr ci = (cptr)(size_t)(const void*)P.iwd;
  const Slots sl(wave, lane);
  const int hi3 = lane | (wave << 6);

  // Layout S: lane l, register k holds index r + 2 l + 128 k + 2048 wave.
  double so[2][16];
  {
    const int vs = ((lane << 1) | (wave << 11)) * 8;
#pragma unroll
    for (int k = 0; k < 16; ++k) io.load(vs, k, so[0][k], so[1][k]);
  }
  // Twiddles: the middle pass's block (into LDS later) and the first
  // pass's runs (bits 3 and 2 whole, the first twiddle of bits 1 and 0).
  const double t2a = __builtin_bit_cast(double, ld64(twr, (kT2Base + t) * 8, 0));
  const double t2b =
      t < kTwWords - 512 ? __builtin_bit_cast(double, ld64(twr, (kT2Base + 512 + t) * 8, 0)) : 0.0;
  double tw3[2], tw2[4];
  tw_run<2>(twr, n - n / 8 + (hi3 << 1), tw3);
  tw_run<4>(twr, n - n / 4 + (hi3 << 2), tw2);
  const double b1 = __builtin_bit_cast(double, ld64(twr, (n - n / 2 + (hi3 << 3)) * 8, 0));
  const double b0 = __builtin_bit_cast(double, ld64(twr, (hi3 << 4) * 8, 0));
  __builtin_amdgcn_sched_barrier(0);
  // Exchange 3 in reverse (inside the wave), round r: index bit 0 = r.
  double z[32];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int k = 0; k < 16; ++k) lds[sl.x3_s + 68 * k] = so[r][k];
    wave_sync_lds();
#pragma unroll
    for (int e = r; e < 32; e += 2) z[e] = lds[sl.x2_l3 + (e >> 1)];
    pin_words(z);
    wave_sync_lds();
  }
  // Bits 0-3 on layout 3 (register bits 0..3).
  {
    double v[16];
#pragma unroll
    for (int c = 1; c < 16; ++c) v[c] = ci[n - 3 * (1 << (31 - __builtin_clz(c))) + c];
    double tw[16];
    tw_from_base<16>(b0, v, tw, ar);
    gs_stage<0>(z, tw, ar);
    inv_reduce_after<A, 0>(z, ar);
    tw_from_base<8>(b1, v, tw, ar);
    gs_stage<1>(z, tw, ar);
    inv_reduce_after<A, 1>(z, ar);
  }
  gs_stage<2>(z, tw2, ar);
  inv_reduce_after<A, 2>(z, ar);
  gs_stage<3>(z, tw3, ar);
  inv_reduce_after<A, 3>(z, ar);
  // The middle pass's twiddles go to LDS; the barrier publishes them.
  t2[t] = t2a;
  if (t < kTwWords - 512) t2[512 + t] = t2b;
  __syncthreads();
  // Exchange 2 in reverse (inside the wave), round r: index bit 4 = r.
  double y[32];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[sl.x2_l3 + e] = z[16 * r + e];
    wave_sync_lds();
#pragma unroll
    for (int e = r; e < 32; e += 2) y[e] = lds[sl.x2_l2 + 17 * (e >> 1)];
    if (r == 0) {
      pin_words(y);
      wave_sync_lds();
    }
  }
  // Bits 4-8 on layout 2 (register bits 0..4), twiddles from LDS.
  {
    const int hi = (lane >> 4) | (wave << 2);
    auto stage = [&](auto bc) {
      constexpr int b = decltype(bc)::value;
      constexpr int kb = b - 4;
      double tw[1 << (4 - kb)];
      const double* src = t2 + (n - (n >> b) - kT2Base) + (hi << (4 - kb));
#pragma unroll
      for (int c = 0; c < (1 << (4 - kb)); ++c) tw[c] = src[c];
      gs_stage<kb>(y, tw, ar);
      inv_reduce_after<A, b>(y, ar);
    };
    stage(std::integral_constant<int, 4>{});
    stage(std::integral_constant<int, 5>{});
    stage(std::integral_constant<int, 6>{});
    stage(std::integral_constant<int, 7>{});
    stage(std::integral_constant<int, 8>{});
  }
  // Exchange 1 in reverse, round r: waves 4 r .. 4 r + 3 write all their
  // words (index bit 13 = r), every wave reads its registers 16 r .. 16 r + 15.
  // The first barrier orders the writes after every wave's exchange-2 reads.
  double x[32];
  __syncthreads();
  if (wave < 4) {
#pragma unroll
    for (int e = 0; e < 32; ++e) lds[sl.x1_l2 + 17 * e] = y[e];
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = lds[sl.x1_l1 + 544 * e];
  pin_words(x);
  __syncthreads();
  if (wave >= 4) {
#pragma unroll
    for (int e = 0; e < 32; ++e) lds[sl.x1_l2 + 17 * e] = y[e];
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 16; ++e) x[16 + e] = lds[sl.x1_l1 + 544 * e];
  // Bits 9-13 on layout 1 (register bits 0..4), twiddles uniform.
  {
    auto stage = [&](auto bc) {
      constexpr int b = decltype(bc)::value;
      constexpr int kb = b - 9;
      double tw[1 << (4 - kb)];
#pragma unroll
      for (int c = 0; c < (1 << (4 - kb)); ++c) tw[c] = ci[n - (n >> b) + c];
      gs_stage<kb>(x, tw, ar);
      inv_reduce_after<A, b>(x, ar);
    };
    stage(std::integral_constant<int, 9>{});
    stage(std::integral_constant<int, 10>{});
    stage(std::integral_constant<int, 11>{});
    stage(std::integral_constant<int, 12>{});
    // The last stage with the scale folded in: u' = (u + v) sc, v' = (u - v)
    // (w sc).  Scale 0 takes sc = 1, whose product is an exact reduction.
    const double w = ci[n - 2];
    const double ws = f64_mulmod_h(w, sc, ar.q, ar.qinv);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const double a = x[e] + x[e + 16];
      const double d = x[e] - x[e + 16];
      x[e] = f64_mulmod_h(a, sc, ar.q, ar.qinv);
      x[e + 16] = f64_mulmod_h(d, ws, ar.q, ar.qinv);
    }
  }
  const int v1 = ((wave << 6) | lane) * 8;
  if constexpr (IO::kAll) {
    io.store_all(v1, x);
  } else {
#pragma unroll
    for (int e = 0; e < 32; ++e) io.store(v1, e, x[e]);
  }
}

}  // namespace w32

// A packed plane row: 16 digits of B bits in W = ceil(16 B / 32) dwords,
// digit i at bits [B i, B i + B) (plane_pack_bits).
template <int B, class F>
__device__ __forceinline__ void pack_row16_fn(const F& digit, uint32_t (&w)[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int o = B * i;
    const uint32_t d = digit(i);
    w[o >> 5] |= d << (o & 31);
    if ((o & 31) + B > 32) w[(o >> 5) + 1] |= d >> (32 - (o & 31));
  }
}
template <int B>
__device__ __forceinline__ void pack_row16(const uint32_t (&dv)[16], uint32_t (&w)[8]) {
  pack_row16_fn<B>([&](int i) { return dv[i]; }, w);
}
template <int B>
__device__ __forceinline__ void unpack_row16(const uint32_t (&w)[8], uint32_t (&dv)[16]) {
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int o = B * i;
    uint32_t v;
    if ((o & 31) + B > 32)
      v = __builtin_amdgcn_alignbit(w[(o >> 5) + 1], w[o >> 5], o & 31);
    else
      v = w[o >> 5] >> (o & 31);
    dv[i] = v & ((1u << B) - 1);
  }
}
}  // namespace shg

#endif  // SHG_NTT_W32_H_
