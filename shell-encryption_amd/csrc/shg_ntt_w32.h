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

}  // namespace w32
}  // namespace shg

#endif  // SHG_NTT_W32_H_
