// Negacyclic NTT / INTT kernels for gfx950 (see shg_ntt_core.h for the
// register/LDS pass structure and the two arithmetic policies).
//
// A slice crosses HBM once per kernel: the first forward pass loads straight
// from HBM (coalesced across lanes), the last forward pass goes through LDS
// for a coalesced store, and the inverse mirrors this.  Limbs with
// n > 2^kMaxLogM (n = 2^15 does not fit in 160 KiB of LDS) first run their
// leading forward stages (or trailing inverse stages) as streaming radix-2
// passes over HBM.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <type_traits>

#include "shg_device.h"
#include "shg_kernels.h"
#include "shg_ntt_core.h"
#include "shg_ntt_wave.h"
#include "shg_ntt_w32.h"

// Per-thread word count (log2) for 2^14-word slices: 16 words with 1024
// threads (four register passes) by default; 32 words with 512 threads
// measured 2% slower forward and 7% slower inverse on MI355X.
#ifndef SHG_LOGE_14
#define SHG_LOGE_14 4
#endif
#ifndef SHG_NTT_DIRECT_STORE
#define SHG_NTT_DIRECT_STORE 1
#endif
// Where the next item's words are loaded with kNttCfgPrefetch: 0 = after
// the first pass's twiddles, 1 = after the last pass's, 2 = right before
// the current item's stores.
#ifndef SHG_NTT_PF_AT
#define SHG_NTT_PF_AT 0
#endif
// The wave-exchange forward kernel loads the next limb into VGPRs at the
// start of the current one (its uniform twiddles come through the scalar
// cache, so only the per-lane twiddles of the last two passes wait behind
// those loads).  2^14 limbs, 8192 per launch: 0.71 ms without either, 0.65
// with scalar twiddles, 0.64 with both; the 16-words-per-thread kernel
// 0.68 ms (one run).
#ifndef SHG_NTT_WAVE_PF
#define SHG_NTT_WAVE_PF 1
#endif
#ifndef SHG_NTT15_HELD
#define SHG_NTT15_HELD 1
#endif
// The key switch's inverse + digit planes at 2^14 (FP64) on the wave layout.
// The key switch's inverse + digit planes at 2^14 (FP64) on the
// two-workgroups-per-CU inverse.
#ifndef SHG_DIG_W32
#define SHG_DIG_W32 1
#endif
#ifndef SHG_DIG_WAVE
#define SHG_DIG_WAVE 1
#endif
// FP64 inverse transforms of 2^13 and 2^14 limbs on the wave-exchange
// kernel by default (kNttCfgWave selects it at 2^12..2^14 either way):
// 0.70 vs 0.75 ms per 8192 limbs of 2^14, 0.67 vs 0.72 at 2^13; at 2^12 the
// two are level (0.62-0.63 ms per 2^25 words).
#ifndef SHG_NTT_INV_WAVE
#define SHG_NTT_INV_WAVE 1
#endif

namespace shg {

// ---------------------------------------------------------------------------
// Kernels.  Work item i = row * 2^HB + h, where row indexes limbs
// (count * num_limbs of them) and h the slice within the limb.  The grid is
// at most the number of resident workgroups; each workgroup loops over items
// i = blockIdx.x, blockIdx.x + gridDim.x, ... and loads the next item's
// words while it transforms the current one.  Vector loads complete in
// issue order, so the prefetch is issued after the first pass's twiddle
// loads: it then overlaps the first pass and its exchange, and the stores
// of the previous item drain meanwhile.  The grid is a multiple of the
// items per limb period (num_limbs * 2^HB) when it can be, so a workgroup
// keeps one prime; with 8 limbs each XCD (workgroup b runs on XCD b mod 8)
// then serves one prime's twiddles from its own L2.
// ---------------------------------------------------------------------------

template <int LOGM, int HB>
__device__ __forceinline__ uint64_t* slice_ptr(uint64_t* data, int item) {
  constexpr int logn = LOGM + HB;
  return data + ((size_t)(item >> HB) << logn) + ((size_t)(item & ((1 << HB) - 1)) << LOGM);
}

// Canonical outputs of a forward run with LAST_TO_LDS = false: thread t
// holds, for each group gamma = gg*T + t of the last pass, the 2^R words
// (gamma << R) .. (gamma << R) + 2^R - 1, stored as 16-byte pairs.
template <int LOGM, int LOGE, class A>
__device__ __forceinline__ void store_fwd_last(uint64_t* g, const typename A::V (&x)[1 << LOGE],
                                               int t, const A& ar) {
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int R = FwdLast<LOGM, LOGE>::R_LAST;
  constexpr int G = FwdLast<LOGM, LOGE>::G_LAST;
  static_assert(R >= 1, "pairs of words");
#pragma unroll
  for (int gg = 0; gg < G; ++gg) {
    u64x2* o = reinterpret_cast<u64x2*>(g + ((size_t)(gg * T + t) << R));
#pragma unroll
    for (int e = 0; e < (1 << R); e += 2) {
      u64x2 v;
      v.x = ar.fwd_out(x[gg * (1 << R) + e]);
      v.y = ar.fwd_out(x[gg * (1 << R) + e + 1]);
      o[e / 2] = v;
    }
  }
}

// HB = log n - LOGM: the number of leading stages already applied by
// streaming passes (limbs larger than one LDS slice).  When the last pass
// covers at most two stages its words are stored straight from VGPRs (two
// 16-byte pieces per group of four words); otherwise they go through LDS.
template <int LOGM, int LOGE, int HB, class A, bool PF>
__global__ void __launch_bounds__((1 << LOGM) >> LOGE)
    ntt_fwd_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                   LimbPrimes lp, int num_limbs, int items) {
  using V = typename A::V;
  constexpr int E = 1 << LOGE;
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr bool kDirectStore =
      SHG_NTT_DIRECT_STORE && FwdLast<LOGM, LOGE>::R_LAST >= 1 && FwdLast<LOGM, LOGE>::R_LAST <= 2;
  constexpr int kPfStage =
      SHG_NTT_PF_AT == 0 ? 0 : (SHG_NTT_PF_AT == 1 ? LOGM - FwdLast<LOGM, LOGE>::R_LAST : -1);
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds = reinterpret_cast<V*>(smem);
  const int t = threadIdx.x;
  int item = blockIdx.x;
  uint64_t pf[E];  // the words of the item to transform next
  if (PF && item < items) {
    const uint64_t* g = slice_ptr<LOGM, HB>(data, item);
#pragma unroll
    for (int e = 0; e < E; ++e) pf[e] = g[e * T + t];
  }
  for (; item < items; item += gridDim.x) {
    const int h = item & ((1 << HB) - 1);
    const A ar(primes[lp.id[(item >> HB) % num_limbs]]);
    uint64_t* g = slice_ptr<LOGM, HB>(data, item);
    // Element e of thread t is word e*T + t of the slice.
    if constexpr (!PF) {
#pragma unroll
      for (int e = 0; e < E; ++e) pf[e] = g[e * T + t];
    }
    V x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = ar.load(pf[e]);
    const int next = item + gridDim.x;
    auto prefetch = [&](auto s0) {
      if constexpr (PF && decltype(s0)::value == kPfStage) {
        if (next < items) {
          const uint64_t* gn = slice_ptr<LOGM, HB>(data, next);
#pragma unroll
          for (int e = 0; e < E; ++e) pf[e] = gn[e * T + t];
        }
      }
    };
    auto prefetch_before_store = [&]() {
      if constexpr (PF && SHG_NTT_PF_AT == 2) {
        if (next < items) {
          const uint64_t* gn = slice_ptr<LOGM, HB>(data, next);
#pragma unroll
          for (int e = 0; e < E; ++e) pf[e] = gn[e * T + t];
        }
      }
    };
    if constexpr (kDirectStore) {
      // The last pass leaves 2^R consecutive words per group in VGPRs: they
      // are stored from there in 16-byte pieces (its reads were followed by
      // a barrier, so the next item may write LDS at once).
      FwdPasses<LOGM, LOGE, 0, A, false, HB>::run(x, lds, lds, t, h, ar, prefetch);
      prefetch_before_store();
      store_fwd_last<LOGM, LOGE>(g, x, t, ar);
    } else {
      // The last pass goes through LDS so that lanes store consecutive words.
      FwdPasses<LOGM, LOGE, 0, A, true, HB>::run(x, lds, lds, t, h, ar, prefetch);
      prefetch_before_store();
      // The next item's first exchange overwrites what is read here: each
      // group of (up to) four reads is pinned before the barrier.
      constexpr int C = E < 4 ? E : 4;
#pragma unroll
      for (int e0 = 0; e0 < E; e0 += C) {
        V y[C];
#pragma unroll
        for (int e = 0; e < C; ++e) y[e] = lds[lpad((e0 + e) * T + t)];
        pin_words(y);
#pragma unroll
        for (int e = 0; e < C; ++e) g[(e0 + e) * T + t] = ar.fwd_out(y[e]);
      }
      __syncthreads();
    }
  }
}

// Forward transform of 2^14-word limbs with each limb split over two
// workgroups: workgroup half h reads the whole limb, computes its half of
// stage 0 (u + w v for h = 0, u - w v for h = 1, pairs (i, i + 2^13)), then
// runs stages 1..13 on that 2^13-word half in 68 KiB of LDS, so two
// workgroups fit on a CU and one's HBM phases overlap the other's
// arithmetic.  Workgroups b and b + 8 (the same XCD under round-robin
// dispatch) take the two halves of one limb, so the second read of the limb
// is served from that XCD's L2.  Persistent: the grid is a multiple of 16 and
// each pair loops over limbs row = slot, slot + grid/2, ...
template <class A>
__global__ void __launch_bounds__(512, 4)
    ntt_fwd_split_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                         LimbPrimes lp, int num_limbs, int rows) {
  using V = typename A::V;
  constexpr int LOGM = 13;
  constexpr int LOGE = 4;
  constexpr int E = 1 << LOGE;
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int H = 1 << LOGM;
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds = reinterpret_cast<V*>(smem);
  const int t = threadIdx.x;
  const int k = blockIdx.x >> 3;
  const int h = k & 1;
  const int step = gridDim.x >> 1;
  for (int row = ((k >> 1) << 3) + (blockIdx.x & 7); row < rows; row += step) {
    const A ar(primes[lp.id[row % num_limbs]]);
    uint64_t* g = data + ((size_t)row << (LOGM + 1));
    const auto w = ar.fwd_tw(1);
    V x[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
      x[e] = ar.ct_half(ar.load(g[e * T + t]), ar.load(g[H + e * T + t]), w, h != 0);
    // The last pass is one stage: each thread stores its pairs of words
    // straight from VGPRs, and the pass's own barrier frees LDS.
    FwdPasses<LOGM, LOGE, 0, A, false, 1>::run(x, lds, lds, t, h, ar);
    store_fwd_last<LOGM, LOGE>(g + h * H, x, t, ar);
  }
}

// Forward transform of 2^(10+LW)-word limbs with one workgroup-wide LDS
// exchange per limb (shg_ntt_wave.h), FP64 arithmetic, persistent over
// limbs like ntt_fwd_kernel.
template <int LW, class A>
__global__ void __launch_bounds__(64 << LW)
    ntt_fwd_wave_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                        LimbPrimes lp, int num_limbs, int items) {
  using Lay = WaveLayout<LW>;
  constexpr int LOGM = Lay::LOGM;
  extern __shared__ __align__(16) unsigned char smem[];
  double* lds = reinterpret_cast<double*>(smem);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  using T1 = WavePassTw<LW, 1, 3, 0>;
  using T2 = WavePassTw<LW, 2, LW - 1, 0>;
  using T3 = WavePassTw<LW, 3, 3, 0>;
  using T4 = WavePassTw<LW, 4, 1, 0>;
  uint64_t pf[16];  // with SHG_NTT_WAVE_PF: the words of the next limb
  if (SHG_NTT_WAVE_PF && (int)blockIdx.x < items) {
    const int tp1 = Lay::thread_part(1, lane, wave);
    const uint64_t* g0 = data + ((size_t)blockIdx.x << LOGM);
#pragma unroll
    for (int e = 0; e < 16; ++e) pf[e] = g0[tp1 + Lay::reg_part(1, e)];
  }
  for (int item = blockIdx.x; item < items; item += gridDim.x) {
    // An opaque copy of the lane index per limb keeps the per-thread
    // twiddle and LDS offsets inside the loop: hoisted out of it they would
    // all stay live across the loop and spill.
    int lane_o = lane;
    asm volatile("" : "+v"(lane_o));
    const A ar(primes[lp.id[item % num_limbs]]);
    uint64_t* g = data + ((size_t)item << LOGM);
    const int tp1 = Lay::thread_part(1, lane_o, wave);
    double tw1[T1::N];
    wave_pass_tw<LW, 1, 3, 0, true>(tw1, 0, 0, ar);
    double x[16];
    if constexpr (SHG_NTT_WAVE_PF) {
#pragma unroll
      for (int e = 0; e < 16; ++e) x[e] = ar.load(pf[e]);
      const int next = item + gridDim.x;
      if (next < items) {
        const uint64_t* gn = data + ((size_t)next << LOGM);
#pragma unroll
        for (int e = 0; e < 16; ++e) pf[e] = gn[tp1 + Lay::reg_part(1, e)];
      }
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) x[e] = ar.load(g[tp1 + Lay::reg_part(1, e)]);
    }
    wave_pass<LW, 1, 3, 0>(x, tw1, ar);
    // Each pass's twiddles are issued before the exchange that precedes it
    // (and not earlier: the scheduling barriers keep one pass's twiddles
    // live at a time).
    __builtin_amdgcn_sched_barrier(0);
    double tw2[T2::N];
    wave_pass_tw<LW, 2, LW - 1, 0, true>(tw2, wave << 10, 0, ar);
    // Every wave has finished the previous limb's in-wave exchanges.
    __syncthreads();
    wave_to_lds<LW, 1>(x, lds, tp1);
    __syncthreads();
    const int tp2 = Lay::thread_part(2, lane_o, wave);
    wave_from_lds<LW, 2>(x, lds, tp2);
    wave_pass<LW, 2, LW - 1, 0>(x, tw2, ar);
    // From here on the wave only touches the slots of its own words.
    const int tp3 = Lay::thread_part(3, lane_o, wave);
    __builtin_amdgcn_sched_barrier(0);
    double tw3[T3::N];
    wave_pass_tw<LW, 3, 3, 0>(tw3, tp3, 0, ar);
    wave_to_lds<LW, 2>(x, lds, tp2);
    wave_sync_lds();
    wave_from_lds<LW, 3>(x, lds, tp3);
    wave_pass<LW, 3, 3, 0>(x, tw3, ar);
    const int tp4 = Lay::thread_part(4, lane_o, wave);
    __builtin_amdgcn_sched_barrier(0);
    double tw4[T4::N];
    wave_pass_tw<LW, 4, 1, 0>(tw4, tp4, 0, ar);
    wave_to_lds<LW, 3>(x, lds, tp3);
    wave_sync_lds();
    wave_from_lds<LW, 4>(x, lds, tp4);
    // The next limb's exchange writes these slots after its first barrier.
    pin_words(x);
    wave_pass<LW, 4, 1, 0>(x, tw4, ar);
    // Registers (e & 3) hold consecutive words; e >> 2 selects index bits 8, 9.
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      u64x2* o = reinterpret_cast<u64x2*>(g + tp4 + Lay::reg_part(4, gg * 4));
#pragma unroll
      for (int e2 = 0; e2 < 4; e2 += 2) {
        u64x2 v;
        v.x = ar.fwd_out(x[gg * 4 + e2]);
        v.y = ar.fwd_out(x[gg * 4 + e2 + 1]);
        o[e2 / 2] = v;
      }
    }
  }
}

// Inverse transform on the wave-exchange layout, FP64: the forward wave
// kernel's passes in the opposite order.  Words are loaded in pass 4's
// layout (index bits 0 and 1 in registers, so each register group is four
// consecutive words read with two 16-byte loads), two exchanges inside each
// wave lead to pass 2's layout and one workgroup-wide exchange to pass 1's,
// whose words are scaled and stored coalesced.  The next limb is loaded
// while the current one is transformed (from the second pass on).
// Barriers: one before the first in-wave write of a limb (other waves may
// still read the previous limb's workgroup-wide exchange there) and one
// inside that exchange.
template <int LW, class A>
__global__ void __launch_bounds__(64 << LW)
    ntt_inv_wave_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                        LimbPrimes lp, int num_limbs, int items, int scale) {
  using Lay = WaveLayout<LW>;
  constexpr int LOGM = Lay::LOGM;
  extern __shared__ __align__(16) unsigned char smem[];
  double* lds = reinterpret_cast<double*>(smem);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  using T1 = WavePassTw<LW, 4, 1, 0>;
  using T2 = WavePassTw<LW, 3, 3, 0>;
  using T3 = WavePassTw<LW, 2, LW - 1, 0>;
  using T4 = WavePassTw<LW, 1, 3, 0>;
  // Register 4 gg + j holds word tp4 + reg_part(4, 4 gg) + j.
  auto load4 = [&](const uint64_t* g, int tp4, uint64_t (&w)[16]) {
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const u64x2* p = reinterpret_cast<const u64x2*>(g + tp4 + Lay::reg_part(4, gg * 4));
      const u64x2 a = p[0];
      const u64x2 b = p[1];
      w[gg * 4] = a.x;
      w[gg * 4 + 1] = a.y;
      w[gg * 4 + 2] = b.x;
      w[gg * 4 + 3] = b.y;
    }
  };
  uint64_t pf[16];
  if ((int)blockIdx.x < items)
    load4(data + ((size_t)blockIdx.x << LOGM), Lay::thread_part(4, lane, wave), pf);
  for (int item = blockIdx.x; item < items; item += gridDim.x) {
    // Opaque lane copy per limb, as in the forward kernel.
    int lane_o = lane;
    asm volatile("" : "+v"(lane_o));
    const A ar(primes[lp.id[item % num_limbs]]);
    uint64_t* g = data + ((size_t)item << LOGM);
    const int tp4 = Lay::thread_part(4, lane_o, wave);
    double tw1[T1::N];
    wave_pass_tw_inv<LW, 4, 1, 0, false>(tw1, tp4, 0, ar);
    double x[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) x[e] = ar.load(pf[e]);
    wave_pass_inv<LW, 4, 1, 0>(x, tw1, ar);
    __builtin_amdgcn_sched_barrier(0);
    const int tp3 = Lay::thread_part(3, lane_o, wave);
    double tw2[T2::N];
    wave_pass_tw_inv<LW, 3, 3, 0, false>(tw2, tp3, 0, ar);
    // The next limb's loads go out after the second pass's twiddle loads:
    // vector memory operations complete in issue order, so twiddles issued
    // behind a limb's worth of loads would wait for all of them.
    __builtin_amdgcn_sched_barrier(0);
    const int next = item + gridDim.x;
    if (next < items) load4(data + ((size_t)next << LOGM), tp4, pf);
    __builtin_amdgcn_sched_barrier(0);
    // Every wave has finished reading the previous limb's exchange.
    __syncthreads();
    wave_to_lds<LW, 4>(x, lds, tp4);
    wave_sync_lds();
    wave_from_lds<LW, 3>(x, lds, tp3);
    wave_pass_inv<LW, 3, 3, 0>(x, tw2, ar);
    __builtin_amdgcn_sched_barrier(0);
    const int tp2 = Lay::thread_part(2, lane_o, wave);
    double tw3[T3::N];
    wave_pass_tw_inv<LW, 2, LW - 1, 0, true>(tw3, wave << 10, 0, ar);
    wave_to_lds<LW, 3>(x, lds, tp3);
    wave_sync_lds();
    wave_from_lds<LW, 2>(x, lds, tp2);
    wave_pass_inv<LW, 2, LW - 1, 0>(x, tw3, ar);
    __builtin_amdgcn_sched_barrier(0);
    const int tp1 = Lay::thread_part(1, lane_o, wave);
    double tw4[T4::N];
    wave_pass_tw_inv<LW, 1, 3, 0, true>(tw4, 0, 0, ar);
    // Each wave writes only its own words (pass 2's waves own the top bits),
    // which no other wave reads before the barrier.
    wave_to_lds<LW, 2>(x, lds, tp2);
    __syncthreads();
    wave_from_lds<LW, 1>(x, lds, tp1);
    wave_inv_last_pass_store<LW>(x, tw4, scale, ar,
                                 [&](int e, uint64_t v) { g[tp1 + Lay::reg_part(1, e)] = v; });
  }
}

// Transforms of 2^14-word limbs on the 32-words-per-lane layouts
// (shg_ntt_w32.h), FP64 arithmetic: one limb per workgroup, two workgroups
// per CU, in place.
struct W32InPlaceFwd {
  __amdgpu_buffer_rsrc_t g;
  const F64Arith* ar;
  static constexpr bool kAux = false;
  struct Aux {};
  __device__ __forceinline__ Aux aux(int, int) const { return {}; }
  __device__ __forceinline__ double load(int v1, int e) const {
    return ar->load(w32::ld64(g, v1, e * 4096));
  }
  __device__ __forceinline__ void store(int vs, int k, double lo, double hi, Aux) const {
    U64Pair v;
    v.x = ar->fwd_out(lo);
    v.y = ar->fwd_out(hi);
    w32::st128(g, v, vs, k * 1024);
  }
};
struct W32InPlaceInv {
  __amdgpu_buffer_rsrc_t g;
  const F64Arith* ar;
  static constexpr bool kAll = false;
  __device__ __forceinline__ void store_all(int, const double (&)[32]) const {}
  __device__ __forceinline__ void load(int vs, int k, double& lo, double& hi) const {
    const U64Pair v = w32::ld128(g, vs, k * 1024);
    lo = ar->load(v.x);
    hi = ar->load(v.y);
  }
  __device__ __forceinline__ void store(int v1, int e, double x) const {
    w32::st64(g, ar->fwd_out(x), v1, e * 4096);
  }
};

__global__ void __launch_bounds__(w32::kThreads, 4)
    ntt_fwd_w32_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                       LimbPrimes lp, int num_limbs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int item = blockIdx.x;
  const PrimeDev& P = primes[lp.id[item % num_limbs]];
  const F64Arith ar(P);
  const W32InPlaceFwd io{w32::rsrc(data + ((size_t)item << w32::kLogN)), &ar};
  w32::fwd_limb(ar, P, reinterpret_cast<double*>(smem), io);
}

// scale: 0 none, 1 n^-1, 2 n^-1 2^-64 (the inverse followed by ExportInt).
__global__ void __launch_bounds__(w32::kThreads, 4)
    ntt_inv_w32_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                       LimbPrimes lp, int num_limbs, int scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int item = blockIdx.x;
  const PrimeDev& P = primes[lp.id[item % num_limbs]];
  const F64Arith ar(P);
  const W32InPlaceInv io{w32::rsrc(data + ((size_t)item << w32::kLogN)), &ar};
  const double sc = scale == 0 ? 1.0 : scale == 1 ? ar.ninv : ar.ninvx;
  w32::inv_limb(ar, P, reinterpret_cast<double*>(smem), sc, io);
}

// HB = log n - LOGM: the number of trailing stages left to streaming passes.
template <int LOGM, int LOGE, int HB, class A, bool PF>
__global__ void __launch_bounds__((1 << LOGM) >> LOGE)
    ntt_inv_kernel(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                   LimbPrimes lp, int num_limbs, int items, int scale) {
  using V = typename A::V;
  constexpr int E = 1 << LOGE;
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int logn = LOGM + HB;
  constexpr int kPfStage =
      SHG_NTT_PF_AT == 0 ? 0 : (SHG_NTT_PF_AT == 1 ? InvLast<LOGM, LOGE, 0, A>::value : -1);
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds = reinterpret_cast<V*>(smem);
  const int t = threadIdx.x;
  int item = blockIdx.x;
  uint64_t pf[E];
  if (PF && item < items) {
    const uint64_t* g = slice_ptr<LOGM, HB>(data, item);
#pragma unroll
    for (int e = 0; e < E; ++e) pf[e] = g[e * T + t];
  }
  for (; item < items; item += gridDim.x) {
    const int h = item & ((1 << HB) - 1);
    const A ar(primes[lp.id[(item >> HB) % num_limbs]]);
    uint64_t* g = slice_ptr<LOGM, HB>(data, item);
    const int next = item + gridDim.x;
    // Stage the current words in LDS (the first pass reads them back in its
    // own layout); the next item's loads follow the first pass's twiddles.
    if constexpr (!PF) {
#pragma unroll
      for (int e = 0; e < E; ++e) pf[e] = g[e * T + t];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) lds[lpad(e * T + t)] = ar.load(pf[e]);
    __syncthreads();
    auto prefetch = [&](auto i0) {
      if constexpr (PF && decltype(i0)::value == kPfStage) {
        if (next < items) {
          const uint64_t* gn = slice_ptr<LOGM, HB>(data, next);
#pragma unroll
          for (int e = 0; e < E; ++e) pf[e] = gn[e * T + t];
        }
      }
    };
    V x[E];
    InvPasses<LOGM, LOGE, 0, A, false, logn>::run(x, lds, t, h, ar, prefetch);
    if constexpr (PF && SHG_NTT_PF_AT == 2) {
      if (next < items) {
        const uint64_t* gn = slice_ptr<LOGM, HB>(data, next);
#pragma unroll
        for (int e = 0; e < E; ++e) pf[e] = gn[e * T + t];
      }
    }
    // The last pass toggles the top local bits: word (e << I0) | gamma.
    // The scaling mode is uniform: one store loop per mode keeps branches
    // out of the per-word path.
    constexpr int I0 = InvLast<LOGM, LOGE, 0, A>::value;
    constexpr int R = LOGM - I0;
    constexpr int G = E >> R;
    auto store_all = [&](auto mode) {
#pragma unroll
      for (int gg = 0; gg < G; ++gg) {
#pragma unroll
        for (int e = 0; e < (1 << R); ++e)
          g[(e << I0) | (gg * T + t)] = ar.inv_out(x[gg * (1 << R) + e], decltype(mode)::value);
      }
    };
    if (scale == 1)
      store_all(std::integral_constant<int, 1>());
    else if (scale == 2)
      store_all(std::integral_constant<int, 2>());
    else
      store_all(std::integral_constant<int, 0>());
    // The last pass read its exchange before its own barrier, so the next
    // item may write LDS without another one.
  }
}

// Inverse transform of one limb per workgroup followed by gadget digit
// extraction into 16-bit planes (launch_ntt_inv_digits).  After the last
// pass the exported values go through LDS in natural order so that thread
// t can gather the E words of consumer indices [t*E, t*E + E) and store each
// plane's E digits with 16-byte writes.
template <int LOGM, int LOGE, class A>
__global__ void __launch_bounds__((1 << LOGM) >> LOGE)
    ntt_inv_digits_kernel(const uint64_t* __restrict__ in, size_t in_stride,
                          uint16_t* __restrict__ planes, GadgetDesc gd,
                          const PrimeDev* __restrict__ primes, LimbPrimes lp, int num_limbs) {
  using V = typename A::V;
  constexpr int E = 1 << LOGE;
  constexpr int T = (1 << LOGM) >> LOGE;
  constexpr int n = 1 << LOGM;
  constexpr int VC = ks_planes_vc(LOGM);
  constexpr int TC = n / VC;
  static_assert(E == 16, "digit planes are written 16 words per thread");
  extern __shared__ __align__(16) unsigned char smem[];
  V* lds = reinterpret_cast<V*>(smem);
  uint64_t* lds_u = reinterpret_cast<uint64_t*>(smem);
  const int row = blockIdx.x;
  const int ct = row / num_limbs;
  const int l = row - ct * num_limbs;
  const int t = threadIdx.x;
  const A ar(primes[lp.id[l]]);
  const uint64_t* g = in + (size_t)ct * in_stride + ((size_t)l << LOGM);
  if constexpr (LOGM == 14 && std::is_same<A, F64Arith>::value && SHG_DIG_WAVE) {
    // The wave-exchange inverse (the exported scale n^-1 2^-64 folded into
    // its last stage); its words reach LDS in natural order once every
    // thread has read the workgroup-wide exchange.
    using Lay = WaveLayout<4>;
    double x[16], tw4[15];
    uint64_t outv[16];
    const int tp1 = wave_inv_core<0>(0, ar, lds, [&](int p) { return g[p]; }, x, tw4);
    wave_inv_last_pass_store<4>(x, tw4, 2, ar, [&](int e, uint64_t v) { outv[e] = v; });
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds_u[lpad(tp1 + Lay::reg_part(1, e))] = outv[e];
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) lds[lpad(e * T + t)] = ar.load(g[e * T + t]);
    __syncthreads();
    V x[E];
    InvPasses<LOGM, LOGE, 0, A, false, LOGM>::run(x, lds, t, 0, ar);
    // The last pass read its exchange before its barrier: LDS is free.
    constexpr int I0 = InvLast<LOGM, LOGE, 0, A>::value;
    constexpr int R = LOGM - I0;
    constexpr int G = E >> R;
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
#pragma unroll
      for (int e = 0; e < (1 << R); ++e)
        lds_u[lpad((e << I0) | (gg * T + t))] = ar.inv_out(x[gg * (1 << R) + e], 2);
    }
  }
  __syncthreads();
  uint64_t v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int idx = t * E + i;
    // Row t of 16 entries: positions (t >> 6) + 16 (t & 63) + 1024 i at
    // n = 2^14 (plane_row14), t + TC i otherwise.
    const int pos = LOGM == 14 && SHG_KS_N14 ? (t >> 6) + 16 * (t & 63) + 1024 * i
                                              : (idx % VC) * TC + idx / VC;
    v[i] = lds_u[lpad(pos)];
  }
  const int lb = gd.limb_log_b[l];
  const int nd = gd.limb_digits[l];
  const uint32_t mask = (1u << lb) - 1;
  uint16_t* base = planes + ((size_t)ct * gd.k + gd.limb_first[l]) * n + (size_t)t * E;
  const int pb = LOGM == 14 ? plane_pack_bits(gd.plane_bits) : 0;
  for (int dd = 0; dd < nd; ++dd) {
    const int sh = dd * lb;
    if (pb) {
      // Packed rows (plane_pack_bits): this thread's row t.
      uint32_t dv[16], w[8];
#pragma unroll
      for (int i = 0; i < 16; ++i) dv[i] = (uint32_t)(v[i] >> sh) & mask;
      const int wn = pb == 8 ? 4 : (pb == 10 ? 5 : 6);
      if (pb == 8)
        pack_row16<8>(dv, w);
      else if (pb == 10)
        pack_row16<10>(dv, w);
      else
        pack_row16<12>(dv, w);
      uint32_t* p32 = reinterpret_cast<uint32_t*>(base - (size_t)t * E + (size_t)dd * n);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        if (k < wn) p32[k * 1024 + t] = w[k];
      continue;
    }
    uint32_t w[E / 2];
#pragma unroll
    for (int i = 0; i < E; i += 2)
      w[i / 2] = ((uint32_t)(v[i] >> sh) & mask) | (((uint32_t)(v[i + 1] >> sh) & mask) << 16);
    uint4* dst = reinterpret_cast<uint4*>(base + (size_t)dd * n);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

// The same at n = 2^14 with FP64 arithmetic on the two-workgroups-per-CU
// inverse: its output layout gives thread t the positions t + 512 e
// (e = 0..31).  With 32 entries per key-switch thread (VC = 32) those are
// exactly the plane entries 32 t .. 32 t + 31 ((p % 512) * 32 + p / 512);
// with 16 (VC = 16), row r of 16 entries holds positions (r >> 6) +
// 16 (r & 63) + 1024 k (plane_row14), so the even e give row
// 64 (t & 15) + (t >> 4) and the odd e the row 32 above it.  Either way each
// plane takes four 16-byte stores per thread and no further exchange.
struct W32Digits {
  __amdgpu_buffer_rsrc_t g;
  const F64Arith* ar;
  uint16_t* base;  // digit 0's plane at this workgroup's first entry
  int nd, lb, plane_bits;
  double* lds;     // the transform's LDS, free once it has run
  static constexpr bool kAll = true;
  __device__ __forceinline__ void load(int vs, int k, double& lo, double& hi) const {
    const U64Pair v = w32::ld128(g, vs, k * 1024);
    lo = ar->load(v.x);
    hi = ar->load(v.y);
  }
  __device__ __forceinline__ void store(int, int, double) const {}
  // Packed planes (plane_pack_bits, B bits per digit): each thread packs its
  // rows plane_row14(t) (even e) and plane_row14(t + 512) (odd e) into W
  // dwords; batches of planes go through LDS (dword k of row r at
  // (plane, k) * 1088 + r + 4 (r >> 6): the rows a wave writes fall on
  // distinct banks) and leave as 16-byte stores of consecutive dwords, since
  // rows 64 apart sit in one wave.
  template <int B>
  __device__ __forceinline__ void store_packed(const uint64_t (&v)[32], int t, uint32_t mask) const {
    constexpr int W = (16 * B + 31) / 32;
    constexpr int kPer = 1088 * W;                   // LDS dwords per plane
    constexpr int kBatch = (2 * w32::kLdsWords) / kPer;  // planes per batch
    uint32_t* l32 = reinterpret_cast<uint32_t*>(lds);
    const int r0 = plane_row14(t), r1 = plane_row14(t + 512);
    const int s0 = r0 + 4 * (r0 >> 6), s1 = r1 + 4 * (r1 >> 6);
    for (int d0 = 0; d0 < nd; d0 += kBatch) {
      const int nb = nd - d0 < kBatch ? nd - d0 : kBatch;
      __syncthreads();  // the previous batch's (or the transform's) LDS reads are done
      for (int i = 0; i < nb; ++i) {
        const int sh = (d0 + i) * lb;
        uint32_t p0[8], p1[8];
        pack_row16_fn<B>([&](int c) { return (uint32_t)(v[2 * c] >> sh) & mask; }, p0);
        pack_row16_fn<B>([&](int c) { return (uint32_t)(v[2 * c + 1] >> sh) & mask; }, p1);
#pragma unroll
        for (int k = 0; k < W; ++k) {
          l32[(i * W + k) * 1088 + s0] = p0[k];
          l32[(i * W + k) * 1088 + s1] = p1[k];
        }
      }
      __syncthreads();
      // Chunk c (4 dwords) of the batch: plane d0 + c / (256 W), dword
      // 4 (c % (256 W)) of that plane, i.e. sub-array k, rows 4 c' .. 4 c' + 3.
      for (int c = t; c < nb * W * 256; c += 512) {
        const int i = c / (256 * W), cw = c - i * 256 * W;
        const int k = cw >> 8, r = (cw & 255) * 4;
        const uint4 val = *reinterpret_cast<const uint4*>(l32 + (i * W + k) * 1088 + r + 4 * (r >> 6));
        uint32_t* dst = reinterpret_cast<uint32_t*>(base + ((size_t)(d0 + i) << w32::kLogN));
        *reinterpret_cast<uint4*>(dst + k * 1024 + r) = val;
      }
    }
  }
  __device__ __forceinline__ void store_all(int v1, const double (&x)[32]) const {
    uint64_t v[32];
#pragma unroll
    for (int e = 0; e < 32; ++e) v[e] = ar->fwd_out(x[e]);
    constexpr int VC = ks_planes_vc(w32::kLogN);
    static_assert(VC == 32 || VC == 16, "plane layouts of 32 or 16 entries per thread");
    const uint32_t mask = (1u << lb) - 1;
    const int t = v1 / 8;
    if constexpr (VC == 16) {
      switch (plane_pack_bits(plane_bits)) {
        case 8: return store_packed<8>(v, t, mask);
        case 10: return store_packed<10>(v, t, mask);
        case 12: return store_packed<12>(v, t, mask);
        default: break;
      }
    }
    for (int dd = 0; dd < nd; ++dd) {
      const int sh = dd * lb;
      uint16_t* plane = base + ((size_t)dd << w32::kLogN);
      auto digit = [&](int e) { return (uint32_t)(v[e] >> sh) & mask; };
      uint32_t w[16];
      if constexpr (VC == 32) {
        // w[i] holds entries 2i, 2i + 1 of the thread's run: e = 2i, 2i + 1.
#pragma unroll
        for (int i = 0; i < 16; ++i) w[i] = digit(2 * i) | (digit(2 * i + 1) << 16);
        uint4* dst = reinterpret_cast<uint4*>(plane + (size_t)t * 32);
#pragma unroll
        for (int j = 0; j < 4; ++j) dst[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
      } else {
        // Even e: row plane_row14(t), odd e: row plane_row14(t + 512), entry
        // e / 2 of each.
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          w[i] = digit(4 * i) | (digit(4 * i + 2) << 16);
          w[8 + i] = digit(4 * i + 1) | (digit(4 * i + 3) << 16);
        }
        uint4* d0 = reinterpret_cast<uint4*>(plane + (size_t)plane_row14(t) * 16);
        uint4* d1 = reinterpret_cast<uint4*>(plane + (size_t)plane_row14(t + 512) * 16);
        d0[0] = make_uint4(w[0], w[1], w[2], w[3]);
        d0[1] = make_uint4(w[4], w[5], w[6], w[7]);
        d1[0] = make_uint4(w[8], w[9], w[10], w[11]);
        d1[1] = make_uint4(w[12], w[13], w[14], w[15]);
      }
    }
  }
};

__global__ void __launch_bounds__(w32::kThreads, 4)
    ntt_inv_digits_w32_kernel(const uint64_t* __restrict__ in, size_t in_stride,
                              uint16_t* __restrict__ planes, GadgetDesc gd,
                              const PrimeDev* __restrict__ primes, LimbPrimes lp, int num_limbs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int row = blockIdx.x;
  const int ct = row / num_limbs;
  const int l = row - ct * num_limbs;
  const PrimeDev& P = primes[lp.id[l]];
  const F64Arith ar(P);
  const W32Digits io{w32::rsrc(in + (size_t)ct * in_stride + ((size_t)l << w32::kLogN)), &ar,
                     planes + (((size_t)ct * gd.k + gd.limb_first[l]) << w32::kLogN),
                     gd.limb_digits[l], gd.limb_log_b[l], gd.plane_bits,
                     reinterpret_cast<double*>(smem)};
  w32::inv_limb(ar, P, reinterpret_cast<double*>(smem), ar.ninvx, io);
}

template <int LOGM, class A>
static hipError_t launch_inv_digits(const uint64_t* in, size_t in_stride, uint16_t* planes,
                                    const GadgetDesc& gd, const PrimeDev* primes,
                                    const LimbPrimes& lp, int count, int num_limbs,
                                    hipStream_t s) {
  if constexpr (LOGM == 14 && std::is_same<A, F64Arith>::value && SHG_DIG_W32) {
    const size_t shmem = (size_t)w32::kLdsWords * 8;
    set_max_dynamic_shmem((const void*)ntt_inv_digits_w32_kernel, shmem);
    hipLaunchKernelGGL(ntt_inv_digits_w32_kernel, dim3((unsigned)(count * num_limbs)),
                       dim3(w32::kThreads), shmem, s, in, in_stride, planes, gd, primes, lp,
                       num_limbs);
    return hipGetLastError();
  }
  constexpr int LOGE = 4;
  constexpr int T = (1 << LOGM) >> LOGE;
  const size_t shmem = (size_t)lds_words<LOGM>() * 8;
  auto k = ntt_inv_digits_kernel<LOGM, LOGE, A>;
  if (shmem > 65536) set_max_dynamic_shmem((const void*)k, shmem);
  hipLaunchKernelGGL(k, dim3((unsigned)(count * num_limbs)), dim3(T), shmem, s, in, in_stride,
                     planes, gd, primes, lp, num_limbs);
  return hipGetLastError();
}

template <class A>
static hipError_t dispatch_inv_digits(int logn, const uint64_t* in, size_t in_stride,
                                      uint16_t* planes, const GadgetDesc& gd,
                                      const PrimeDev* primes, const LimbPrimes& lp, int count,
                                      int num_limbs, hipStream_t s) {
#define SHG_DIG(M) launch_inv_digits<M, A>(in, in_stride, planes, gd, primes, lp, count, num_limbs, s)
  switch (logn) {
    case 4: return SHG_DIG(4);
    case 5: return SHG_DIG(5);
    case 6: return SHG_DIG(6);
    case 7: return SHG_DIG(7);
    case 8: return SHG_DIG(8);
    case 9: return SHG_DIG(9);
    case 10: return SHG_DIG(10);
    case 11: return SHG_DIG(11);
    case 12: return SHG_DIG(12);
    case 13: return SHG_DIG(13);
    case 14: return SHG_DIG(14);
    default: return hipErrorInvalidValue;
  }
#undef SHG_DIG
}

hipError_t launch_ntt_inv_digits(int variant, const uint64_t* in, size_t in_stride,
                                 uint16_t* planes, const GadgetDesc& gd, const PrimeDev* primes,
                                 const LimbPrimes& lp, int count, int num_limbs, int logn,
                                 hipStream_t s) {
  if ((variant & 3) == SHG_NTT_F64)
    return dispatch_inv_digits<F64Arith>(logn, in, in_stride, planes, gd, primes, lp, count,
                                         num_limbs, s);
  return dispatch_inv_digits<IntArith>(logn, in, in_stride, planes, gd, primes, lp, count,
                                       num_limbs, s);
}

// Streaming radix-2 stage over HBM for limbs larger than the LDS slice:
// forward stage s (pairs at distance n/2^(s+1)), or inverse stage i.
// Integer values are kept in [0, 2q); F64 values are kept balanced.
template <class A, bool INVERSE>
__global__ void ntt_stream_stage(uint64_t* __restrict__ data, const PrimeDev* __restrict__ primes,
                                 LimbPrimes lp, int num_limbs, int logn, int stage,
                                 long long total_pairs) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total_pairs) return;
  const int half_pairs_log = logn - 1;
  const long long row = gid >> half_pairs_log;
  const int p = (int)(gid & ((1 << half_pairs_log) - 1));
  const PrimeDev& P = primes[lp.id[row % num_limbs]];
  const A ar(P);
  uint64_t* g = data + ((size_t)row << logn);
  const int bit = INVERSE ? stage : (logn - 1 - stage);
  const int lo = p & ((1 << bit) - 1);
  const int hi = p >> bit;
  const int i0 = (hi << (bit + 1)) | lo;
  const int i1 = i0 | (1 << bit);
  typename A::V u = ar.load(g[i0]);
  typename A::V v = ar.load(g[i1]);
  if (!INVERSE) {
    const auto w = ar.fwd_tw((1 << stage) + hi);
    ar.ct(u, v, w);
    g[i0] = ar.fwd_out(u);
    g[i1] = ar.fwd_out(v);
  } else {
    const int n = 1 << logn;
    const auto w = ar.inv_tw(n - (n >> stage) + hi);
    ar.gs(u, v, w);
    g[i0] = ar.fwd_out(u);
    g[i1] = ar.fwd_out(v);
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

// Raises a kernel's dynamic LDS limit once per (kernel, device): later
// launches (for instance under graph capture) make no attribute calls.
void set_max_dynamic_shmem(const void* k, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> done;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  size_t& cur = done[{k, dev}];
  if (cur >= bytes) return;
  if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess)
    cur = bytes;
}

// Prefetching the next item (kNttCfgPrefetch) costs 2E VGPRs.  At E = 16
// (1024 threads, 128 VGPRs) it spills and measured slower (fwd 0.80 vs 0.63
// ms per 8192 limbs of 2^14); the autotuner may still pick it for a shape.

// Workgroups resident at once on the current device for kernel k (at least
// one per CU), capped by the item count and rounded down to a multiple of
// `period` when that leaves at least one period.
// The residency per (kernel, device) is cached, so launches after the first
// (for instance while a graph is being captured) make no runtime queries.
static int persistent_grid(const void* k, int threads, size_t shmem, int items, int period) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, long long> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return items;
  long long cap = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({k, dev});
    if (it != cache.end()) cap = it->second;
  }
  if (cap == 0) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      return items;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, shmem) != hipSuccess ||
        per_cu <= 0)
      per_cu = 1;
    cap = (long long)cus * per_cu;
    std::lock_guard<std::mutex> lock(mu);
    cache[{k, dev}] = cap;
  }
  if (cap >= items) return items;
  if (period > 0 && cap >= period) cap -= cap % period;
  return (int)cap;
}

template <int LOGM, int LOGE, int HB, class A, bool PF>
static hipError_t launch_fused_hb(bool inverse, uint64_t* data, const PrimeDev* primes,
                                  const LimbPrimes& lp, int rows, int num_limbs, int scale,
                                  hipStream_t stream) {
  constexpr int T = (1 << LOGM) >> LOGE;
  const size_t shmem = (size_t)lds_words<LOGM>() * 8;
  const int items = rows << HB;
  const int period = num_limbs << HB;
  if (!inverse) {
    auto k = ntt_fwd_kernel<LOGM, LOGE, HB, A, PF>;
    if (shmem > 65536)
      set_max_dynamic_shmem((const void*)k, shmem);
    const int grid = persistent_grid((const void*)k, T, shmem, items, period);
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), shmem, stream, data, primes, lp, num_limbs, items);
  } else {
    auto k = ntt_inv_kernel<LOGM, LOGE, HB, A, PF>;
    if (shmem > 65536)
      set_max_dynamic_shmem((const void*)k, shmem);
    const int grid = persistent_grid((const void*)k, T, shmem, items, period);
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), shmem, stream, data, primes, lp, num_limbs, items,
                       scale);
  }
  return hipGetLastError();
}

template <class A>
static hipError_t launch_fwd_split(uint64_t* data, const PrimeDev* primes, const LimbPrimes& lp,
                                   int rows, int num_limbs, hipStream_t stream) {
  constexpr int T = 512;
  const size_t shmem = (size_t)lds_words<13>() * 8;
  auto k = ntt_fwd_split_kernel<A>;
  set_max_dynamic_shmem((const void*)k, shmem);
  // Two workgroups per limb; the grid must be a multiple of 16.
  int grid = persistent_grid((const void*)k, T, shmem, 2 * rows, 0);
  grid = grid < 16 ? 16 : grid - grid % 16;
  hipLaunchKernelGGL(k, dim3(grid), dim3(T), shmem, stream, data, primes, lp, num_limbs, rows);
  return hipGetLastError();
}

template <int LW>
static hipError_t launch_fwd_wave(uint64_t* data, const PrimeDev* primes, const LimbPrimes& lp,
                                  int rows, int num_limbs, hipStream_t stream) {
  constexpr int T = 64 << LW;
  const size_t shmem = (size_t)lds_words<10 + LW>() * 8;
  auto k = ntt_fwd_wave_kernel<LW, F64Arith>;
  if (shmem > 65536) set_max_dynamic_shmem((const void*)k, shmem);
  const int grid = persistent_grid((const void*)k, T, shmem, rows, num_limbs);
  hipLaunchKernelGGL(k, dim3(grid), dim3(T), shmem, stream, data, primes, lp, num_limbs, rows);
  return hipGetLastError();
}

template <int LW>
static hipError_t launch_inv_wave(uint64_t* data, const PrimeDev* primes, const LimbPrimes& lp,
                                  int rows, int num_limbs, int scale, hipStream_t stream) {
  constexpr int T = 64 << LW;
  const size_t shmem = (size_t)lds_words<10 + LW>() * 8;
  auto k = ntt_inv_wave_kernel<LW, F64Arith>;
  if (shmem > 65536) set_max_dynamic_shmem((const void*)k, shmem);
  const int grid = persistent_grid((const void*)k, T, shmem, rows, num_limbs);
  hipLaunchKernelGGL(k, dim3(grid), dim3(T), shmem, stream, data, primes, lp, num_limbs, rows,
                     scale);
  return hipGetLastError();
}

static hipError_t launch_w32(bool inverse, uint64_t* data, const PrimeDev* primes,
                             const LimbPrimes& lp, int rows, int num_limbs, int scale,
                             hipStream_t stream) {
  const size_t shmem = (size_t)w32::kLdsWords * 8;
  if (!inverse) {
    auto k = ntt_fwd_w32_kernel;
    set_max_dynamic_shmem((const void*)k, shmem);
    hipLaunchKernelGGL(k, dim3(rows), dim3(w32::kThreads), shmem, stream, data, primes, lp,
                       num_limbs);
  } else {
    auto k = ntt_inv_w32_kernel;
    set_max_dynamic_shmem((const void*)k, shmem);
    hipLaunchKernelGGL(k, dim3(rows), dim3(w32::kThreads), shmem, stream, data, primes, lp,
                       num_limbs, scale);
  }
  return hipGetLastError();
}

template <int LOGM, int LOGE, class A, bool PF>
static hipError_t launch_fused(bool inverse, uint64_t* data, const PrimeDev* primes,
                               const LimbPrimes& lp, int rows, int num_limbs, int logn,
                               int scale, hipStream_t stream) {
  if (logn == LOGM)
    return launch_fused_hb<LOGM, LOGE, 0, A, PF>(inverse, data, primes, lp, rows, num_limbs,
                                                 scale, stream);
  // Half-limb slices: 2^15 limbs on 2^14 slices, and the stage split of
  // 2^13 and 2^14 limbs (without prefetch).
  if constexpr (LOGM == kMaxLogM || (LOGM >= 12 && !PF)) {
    if (logn == LOGM + 1)
      return launch_fused_hb<LOGM, LOGE, 1, A, PF>(inverse, data, primes, lp, rows, num_limbs,
                                                   scale, stream);
  }
  return hipErrorInvalidValue;
}

template <class A>
static hipError_t launch_stream(bool inverse, uint64_t* data, const PrimeDev* primes,
                                const LimbPrimes& lp, int rows, int num_limbs, int logn,
                                int stage, hipStream_t stream) {
  const long long pairs = (long long)rows << (logn - 1);
  const int block = 256;
  const unsigned grid = (unsigned)((pairs + block - 1) / block);
  if (!inverse)
    hipLaunchKernelGGL((ntt_stream_stage<A, false>), dim3(grid), dim3(block), 0, stream, data,
                       primes, lp, num_limbs, logn, stage, pairs);
  else
    hipLaunchKernelGGL((ntt_stream_stage<A, true>), dim3(grid), dim3(block), 0, stream, data,
                       primes, lp, num_limbs, logn, stage, pairs);
  return hipGetLastError();
}

// Slices of 2^12..2^14 words have two register tilings: the default (E =
// 16 words per thread) and the alternate one selected by kNttCfgAlt (E = 8
// below 2^14, E = 32 at 2^14).
template <class A, bool PF>
static hipError_t dispatch_fused(int logm, bool alt, bool inverse, uint64_t* data,
                                 const PrimeDev* primes, const LimbPrimes& lp, int rows,
                                 int num_limbs, int logn, int scale, hipStream_t s) {
#define SHG_FUSED(M, E) launch_fused<M, E, A, PF>(inverse, data, primes, lp, rows, num_limbs, logn, scale, s)
  switch (logm) {
    case 1: return SHG_FUSED(1, 1);
    case 2: return SHG_FUSED(2, 2);
    case 3: return SHG_FUSED(3, 3);
    case 4: return SHG_FUSED(4, 4);
    case 5: return SHG_FUSED(5, 4);
    case 6: return SHG_FUSED(6, 4);
    case 7: return SHG_FUSED(7, 4);
    case 8: return SHG_FUSED(8, 4);
    case 9: return SHG_FUSED(9, 4);
    case 10: return SHG_FUSED(10, 4);
    case 11: return SHG_FUSED(11, 4);
    case 12: return alt ? SHG_FUSED(12, 3) : SHG_FUSED(12, 4);
    case 13: return alt ? SHG_FUSED(13, 3) : SHG_FUSED(13, 4);
    case 14: return alt ? SHG_FUSED(14, 5) : SHG_FUSED(14, SHG_LOGE_14);
    default: return hipErrorInvalidValue;
  }
#undef SHG_FUSED
}

template <class A, bool PF>
static hipError_t run_ntt(bool inverse, bool alt, uint64_t* data, const PrimeDev* primes,
                          const LimbPrimes& lp, int rows, int num_limbs, int logn, int max_logm,
                          int inv_mode, hipStream_t s) {
  const int logm = logn < max_logm ? logn : max_logm;
  hipError_t e = hipSuccess;
  if (!inverse) {
    for (int st = 0; st < logn - logm && e == hipSuccess; ++st)
      e = launch_stream<A>(false, data, primes, lp, rows, num_limbs, logn, st, s);
    if (e == hipSuccess)
      e = dispatch_fused<A, PF>(logm, alt, false, data, primes, lp, rows, num_limbs, logn, 1, s);
  } else {
    e = dispatch_fused<A, PF>(logm, alt, true, data, primes, lp, rows, num_limbs, logn,
                              logn == logm ? inv_mode : 0, s);
    for (int st = logm; st < logn && e == hipSuccess; ++st) {
      e = launch_stream<A>(true, data, primes, lp, rows, num_limbs, logn, st, s);
    }
    if (e == hipSuccess && logn > logm)
      e = launch_scale_ninv(data, primes, lp, rows, num_limbs, logn, inv_mode, s);
  }
  return e;
}

hipError_t launch_ntt(bool inverse, int cfg, uint64_t* data, const PrimeDev* primes,
                      const LimbPrimes& lp, int rows, int num_limbs, int logn, int max_logm,
                      hipStream_t s, int inv_mode) {
  const bool alt = (cfg & kNttCfgAlt) != 0;
  if (cfg & kNttCfgHalf) {
    // One streamed stage and half-limb slices (no prefetching variant).
    if (logn < 13 || logn > 15 || logn - 1 > max_logm) return hipErrorInvalidValue;
    if ((cfg & 3) == SHG_NTT_F64)
      return run_ntt<F64Arith, false>(inverse, alt, data, primes, lp, rows, num_limbs, logn,
                                      logn - 1, inv_mode, s);
    return run_ntt<IntArith, false>(inverse, alt, data, primes, lp, rows, num_limbs, logn,
                                    logn - 1, inv_mode, s);
  }
  // FP64 transforms of 2^14 limbs default to the kernels with two
  // workgroups per CU (forward 0.47 vs 0.58 ms, inverse 0.50 vs 0.62 ms per
  // 8192 limbs against the wave-exchange kernels on one box); the other
  // flags select the other kernels.
  const bool pair_default =
      logn == 14 && (cfg & (kNttCfgAlt | kNttCfgPrefetch | kNttCfgSplit | kNttCfgWave)) == 0;
  if (((cfg & kNttCfgPair) || pair_default) && (cfg & 3) == SHG_NTT_F64 && logn == 14 &&
      max_logm >= 14)
    return launch_w32(inverse, data, primes, lp, rows, num_limbs, inverse ? inv_mode : 1, s);
  // FP64 forward transforms of 2^12..2^14 limbs with the wave flag take the
  // wave-exchange kernel; so do inverse ones, and 2^13 inverses by default.
  if (!inverse && (cfg & kNttCfgWave) && (cfg & 3) == SHG_NTT_F64 &&
      logn >= 12 && logn <= 14 && logn <= max_logm) {
    if (logn == 12) return launch_fwd_wave<2>(data, primes, lp, rows, num_limbs, s);
    if (logn == 13) return launch_fwd_wave<3>(data, primes, lp, rows, num_limbs, s);
    return launch_fwd_wave<4>(data, primes, lp, rows, num_limbs, s);
  }
  const bool inv_wave_default = SHG_NTT_INV_WAVE && logn == 13 &&
                                (cfg & (kNttCfgAlt | kNttCfgPrefetch | kNttCfgSplit)) == 0;
  if (inverse && ((cfg & kNttCfgWave) || inv_wave_default) &&
      (cfg & 3) == SHG_NTT_F64 && logn >= 12 && logn <= 14 && logn <= max_logm) {
    if (logn == 12) return launch_inv_wave<2>(data, primes, lp, rows, num_limbs, inv_mode, s);
    if (logn == 13) return launch_inv_wave<3>(data, primes, lp, rows, num_limbs, inv_mode, s);
    return launch_inv_wave<4>(data, primes, lp, rows, num_limbs, inv_mode, s);
  }
  // 2^15 limbs: one kernel with both halves held in VGPRs (instead of a
  // streaming stage over HBM followed by the 2^14-slice kernel).
  if (SHG_NTT15_HELD && logn == 15 && max_logm >= 14 && (cfg & 3) == SHG_NTT_F64 &&
      (cfg & (kNttCfgAlt | kNttCfgPrefetch | kNttCfgSplit | kNttCfgWave)) == 0 &&
      (!inverse || inv_mode == 1))
    return launch_ntt15(inverse, data, primes, lp, rows, num_limbs, s);
  if (!inverse && logn == 14 && max_logm >= 13 && (cfg & kNttCfgSplit)) {
    if ((cfg & 3) == SHG_NTT_F64)
      return launch_fwd_split<F64Arith>(data, primes, lp, rows, num_limbs, s);
    return launch_fwd_split<IntArith>(data, primes, lp, rows, num_limbs, s);
  }
  if ((cfg & 3) == SHG_NTT_F64) {
    if (cfg & kNttCfgPrefetch)
      return run_ntt<F64Arith, true>(inverse, alt, data, primes, lp, rows, num_limbs, logn,
                                     max_logm, inv_mode, s);
    return run_ntt<F64Arith, false>(inverse, alt, data, primes, lp, rows, num_limbs, logn,
                                    max_logm, inv_mode, s);
  }
  return run_ntt<IntArith, false>(inverse, alt, data, primes, lp, rows, num_limbs, logn, max_logm,
                                  inv_mode, s);
}

}  // namespace shg
