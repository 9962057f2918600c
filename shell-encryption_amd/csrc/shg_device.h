// Device-side data structures and modular arithmetic for the gfx950 kernels.
//
// Residues are MontgomeryInt<Uint64> representations in [0, q), q < 2^62
// (q < 2^50 for the FP64 transform variant).  Every routine here computes an
// exact modular result, so any of them may stand in for the reference's
// arithmetic as long as the final value is reduced to the same canonical
// representative (shell_encryption/montgomery.h:318-465).
#ifndef SHG_DEVICE_H_
#define SHG_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shell_gpu.h"

namespace shg {

// Per-(q, log_n) constants, resident in device memory.  Twiddle tables are
// the reference's psis_bitrev_constant / psis_inv_bitrev_constant
// (ntt_parameters.h:160-179) interleaved as (c, cb) pairs, plus FP64 copies
// of the values c.
struct PrimeDev {
  uint64_t q;
  uint64_t inv_modulus;    // q * inv_modulus = -1 mod 2^64 (REDC)
  uint64_t r_mod_q;        // Montgomery form of 1
  uint64_t r_barrett;      // floor(r_mod_q * 2^64 / q): ImportInt
  uint64_t inv_r;          // 2^-64 mod q: ExportInt
  uint64_t inv_r_barrett;  // floor(inv_r * 2^64 / q)
  uint64_t bn;             // floor(2^64 / q): Barrett reduction
  uint64_t n_inv_c;        // value of n^-1 (GetConstant().first)
  uint64_t n_inv_cb;       // its Shoup constant
  double qd;               // q as double
  double qinv;             // fl(1/q)
  double n_inv_d;          // value of n^-1 as double
  double n_inv_dq;         // fl(n^-1 / q)
  uint64_t n_inv_x_c;      // value of n^-1 * 2^-64 (INTT that also exports)
  uint64_t n_inv_x_cb;     // its Shoup constant
  double n_inv_x_d;        // n^-1 * 2^-64 as double
  double n_inv_x_dq;       // fl(n^-1 * 2^-64 / q)
  int log_n;
  int log_q;
  const uint64_t* fw;      // [2n] forward (c, cb), bit-reversed psi powers
  const uint64_t* iw;      // [2n] inverse (c, cb)
  const double* fwd;       // [n] forward twiddle values as doubles
  const double* iwd;       // [n] inverse twiddle values as doubles
};

// Pointer into global memory.  Table pointers read from a PrimeDev would
// otherwise be generic (flat) pointers, whose loads also count against the
// LDS wait counter and so serialise with the LDS exchanges.
template <class T>
using gptr = const T __attribute__((address_space(1)))*;
// A table of doubles read through buffer loads: the descriptor lives in
// SGPRs and each lane supplies a 32-bit byte offset, so an indexed read
// costs one shift where a 64-bit address costs three VALU operations per
// lane.  Descriptor word 3 (0x00020000) selects the raw 32-bit data format
// of gfx9-family targets; num_records is not used as a bound (callers index
// within the table).
struct F64Table {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ explicit F64Table(const void* p)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff,
                                            0x00020000)) {}
  __device__ __forceinline__ double operator[](uint32_t i) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(i * 8u), 0, 0));
  }
};

// Writable global pointer: data pointers read from device tables (rather
// than kernel arguments) are otherwise flat, and flat stores may alias LDS,
// which serialises them with the LDS exchanges.
template <class T>
using gmut = T __attribute__((address_space(1)))*;

struct LimbPrimes {
  int id[SHG_MAX_LIMBS];
};

// ---------------------------------------------------------------------------
// 64-bit integer arithmetic from 32-bit VALU multiplies.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
  return __umul64hi(a, b);
}

// Shoup product x * c mod q with cb = floor(c * 2^64 / q): the result is
// congruent to x*c and lies in [0, 2q) for any 64-bit x.
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t x, uint64_t c, uint64_t cb, uint64_t q) {
  return x * c - mulhi64(x, cb) * q;
}

__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t m) { return x >= m ? x - m : x; }

// x mod q for any 64-bit x (Barrett with bn = floor(2^64/q)).
__device__ __forceinline__ uint64_t reduce_full(uint64_t x, uint64_t q, uint64_t bn) {
  return csub(x - mulhi64(x, bn) * q, q);
}

// REDC(a*b) = a*b*2^-64 mod q, canonical (montgomery.h:318-336).
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, const PrimeDev& P) {
  const uint64_t lo = a * b;
  const uint64_t hi = mulhi64(a, b);
  const uint64_t u = lo * P.inv_modulus;
  // (a*b + u*q) / 2^64; the low word sums to 0 with a carry iff lo != 0.
  const uint64_t t = hi + mulhi64(u, P.q) + (lo != 0 ? 1 : 0);
  return csub(t, P.q);
}

// this + a*b*2^-64 mod q, canonical (montgomery.h:446-465).
__device__ __forceinline__ uint64_t mont_fma(uint64_t acc, uint64_t a, uint64_t b,
                                             const PrimeDev& P) {
  return csub(acc + mont_mul(a, b, P), P.q);
}

// ImportInt: n * 2^64 mod q for any 64-bit n (montgomery.cc:26-40).
__device__ __forceinline__ uint64_t mont_import(uint64_t n, const PrimeDev& P) {
  return csub(shoup_lazy(n, P.r_mod_q, P.r_barrett, P.q), P.q);
}

// ExportInt: x * 2^-64 mod q (montgomery.h:156-164).
__device__ __forceinline__ uint64_t mont_export(uint64_t x, const PrimeDev& P) {
  return csub(shoup_lazy(x, P.inv_r, P.inv_r_barrett, P.q), P.q);
}

// AddInPlace / SubInPlace for inputs in [0, q] (q itself arises from
// NegateInPlace(0)); the result is canonical as BarrettReduce's is
// (montgomery.h:383-415).
__device__ __forceinline__ uint64_t mod_add(uint64_t a, uint64_t b, uint64_t q) {
  return csub(csub(a + b, q), q);
}
__device__ __forceinline__ uint64_t mod_sub(uint64_t a, uint64_t b, uint64_t q) {
  return csub(csub(a + (q - b), q), q);
}

// ImportInt mod P.q of the integer a value d in [0, qs) stands for under the
// balanced rule of ConvertModulusBalanced (modulus_conversion.h:49-81):
// d > qs/2 is d - qs.  For P.q == qs this is ImportInt(d) itself.
__device__ __forceinline__ uint64_t import_balanced(uint64_t d, uint64_t qs, const PrimeDev& P) {
  if (d > (qs >> 1)) {
    const uint64_t m = mont_import(qs - d, P);
    return m ? P.q - m : 0;
  }
  return mont_import(d, P);
}

// ConvertModulusBalanced (modulus_conversion.h:30-83) from q_l to each q_i.
__device__ __forceinline__ uint64_t convert_balanced(uint64_t rep_q, const PrimeDev& Pq,
                                                     const PrimeDev& Pp) {
  if (Pq.q == Pp.q) return rep_q;
  const uint64_t x = mont_export(rep_q, Pq);
  uint64_t xp = mont_import(x, Pp);
  if (x > (Pq.q >> 1)) {
    const uint64_t diff = Pq.q > Pp.q ? Pq.q - Pp.q : Pp.q - Pq.q;
    const uint64_t diff_p = mont_import(diff, Pp);
    xp = (Pp.q > Pq.q) ? mod_add(xp, diff_p, Pp.q) : mod_sub(xp, diff_p, Pp.q);
  }
  return xp;
}

// ---------------------------------------------------------------------------
// FP64 exact modular arithmetic (q < 2^50).  Integers are held exactly in
// doubles; x*w = h + l exactly (h = fl(x*w), l = fma(x, w, -h)).  The
// quotient c = rint(x * fl(w/q)), or rint(h * fl(1/q)), is within
// 1/2 + 1.5 * 2^-52 |x| w/q of x*w/q (three roundings), so h - c*q is an
// exact small integer and t = (h - c*q) + l = x*w - c*q is exact with
// |t| <= q/2 + 1.5 * 2^-52 |x| w.
// ---------------------------------------------------------------------------

// x < 2^52 -> double, exactly: 2^52 + x has the bits 0x433 << 52 | x.
__device__ __forceinline__ double u52_to_f64(uint64_t x) {
  return __longlong_as_double((long long)(x | 0x4330000000000000ull)) - 4503599627370496.0;
}

// x integral in [0, 2^52) -> uint64: the mantissa of 2^52 + x.
__device__ __forceinline__ uint64_t f64_to_u52(double x) {
  return (uint64_t)__double_as_longlong(x + 4503599627370496.0) & 0x000FFFFFFFFFFFFFull;
}

__device__ __forceinline__ double f64_mulmod(double x, double w, double wq, double q) {
  const double h = x * w;
  const double l = fma(x, w, -h);
  const double c = rint(x * wq);
  const double d = fma(-c, q, h);
  return d + l;
}

// The same product with the quotient taken from h: c = rint(fl(h * fl(1/q))).
// fl(x*w), fl(1/q) and the product are three roundings, as in x * fl(w/q)
// above, so the bound on |t| is the same; it needs no w/q per multiplier,
// which saves one multiply wherever w is not reused.
__device__ __forceinline__ double f64_mulmod_h(double x, double w, double q, double qinv) {
  const double h = x * w;
  const double l = fma(x, w, -h);
  const double c = rint(h * qinv);
  const double d = fma(-c, q, h);
  return d + l;
}

// x - rint(x/q)*q: |result| <= q/2 (+ a negligible margin) for |x| < 2^53.
__device__ __forceinline__ double f64_reduce(double x, double q, double qinv) {
  return fma(-rint(x * qinv), q, x);
}

// Balanced (|x| < q) -> canonical [0, q).
__device__ __forceinline__ double f64_canon(double x, double q) { return x < 0.0 ? x + q : x; }

// f64_to_u52(f64_canon(x, q)) for integral x in (-q, q) with one addition:
// the mantissa of 2^52 + x, or of 2^52 + q + x for x < 0 (q + 2^52 is
// exact for q < 2^52 and loop-invariant).
__device__ __forceinline__ uint64_t f64_canon_u52(double x, double q) {
  const double two52 = 4503599627370496.0;
  const double y = x + (x < 0.0 ? q + two52 : two52);
  return (uint64_t)__double_as_longlong(y) & 0x000FFFFFFFFFFFFFull;
}

}  // namespace shg

#endif  // SHG_DEVICE_H_
