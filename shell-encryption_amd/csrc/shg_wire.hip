// Wire formats of device-resident polynomials and the seeded random pad.
//
// MontgomeryInt::SerializeVector (montgomery.cc:143-167) stores each residue
// of a limb in log2(q) bits through TranscribeBits (transcription.h:60-150):
// the limb becomes one little-endian bit stream, residue j in stream bits
// [j*b, (j+1)*b), packed into bytes low bit first.  RnsPolynomial::Serialize
// (rns/rns_polynomial.h:124-140) keeps one such byte string per limb.  Here a
// polynomial's byte strings sit back to back in device memory; packing and
// unpacking are streaming kernels (one thread per 64-bit stream word, or per
// residue), so keys and database records load from their compact form
// straight into device buffers.
//
// RnsGaloisKey keeps only key_b on the wire and regenerates key_a from a PRNG
// seed (rns/rns_galois_key.cc:81-105, 204-236): RnsPolynomial::SampleUniform
// (rns/rns_polynomial.h:640-667) over SingleThreadChaChaPrng, whose stream is
// ChaCha20 (RFC 7539) under nonce "salt" || LE64(s) for buffer s of 8160
// bytes (prng/chacha_prng_util.cc:27-104).  Each residue takes the low
// log2(q) bits of the next little-endian 64-bit draw, rejected while >= q
// (montgomery.h:260-273, 631-660), so later residues depend on how many
// draws earlier ones rejected.  One workgroup walks the stream: its 1024
// threads each compute one ChaCha block (8 draws), a workgroup-wide prefix
// sum over the accepted draws places them, and the position after the last
// draw a segment (one limb of one polynomial) used starts the next one.
#include <hip/hip_runtime.h>

#include "shg_device.h"
#include "shg_kernels.h"

namespace shg {

namespace {

__device__ __forceinline__ uint64_t low_mask(int b) {
  return b >= 64 ? ~0ull : ((1ull << b) - 1);
}

inline unsigned grid_of(long long items, int block) {
  return (unsigned)((items + block - 1) / block);
}

// Thread per (polynomial, limb, 64-bit word of the limb's stream).
__global__ void pack_kernel(const uint64_t* __restrict__ polys, uint8_t* __restrict__ out,
                            WireDesc wd, long long max_words, long long total) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const long long w = g % max_words;
  const long long pl = g / max_words;
  const int limb = (int)(pl % wd.num_limbs);
  const long long p = pl / wd.num_limbs;
  const int b = wd.bits[limb];
  const long long n = 1LL << wd.log_n;
  const long long limb_bytes = (long long)wd.limb_off[limb + 1] - (long long)wd.limb_off[limb];
  if (w * 8 >= limb_bytes) return;
  const uint64_t* c = polys + (p * wd.num_limbs + limb) * n;
  const uint64_t m = low_mask(b);
  const long long s = w * 64;
  long long j = s / b;
  const int off = (int)(s - j * b);
  uint64_t v = (c[j] & m) >> off;
  int filled = b - off;
  while (filled < 64 && ++j < n) {
    v |= (c[j] & m) << filled;
    filled += b;
  }
  uint8_t* dst = out + p * (long long)wd.limb_off[wd.num_limbs] + wd.limb_off[limb] + w * 8;
  if (wd.aligned) {
    *reinterpret_cast<uint64_t*>(dst) = v;
  } else {
    const long long nb = limb_bytes - w * 8 < 8 ? limb_bytes - w * 8 : 8;
    for (int k = 0; k < nb; ++k) dst[k] = (uint8_t)(v >> (8 * k));
  }
}

// Thread per (polynomial, limb, residue).
__global__ void unpack_kernel(const uint8_t* __restrict__ in, uint64_t* __restrict__ polys,
                              WireDesc wd, long long total) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const long long n = 1LL << wd.log_n;
  const long long j = g & (n - 1);
  const long long pl = g >> wd.log_n;
  const int limb = (int)(pl % wd.num_limbs);
  const long long p = pl / wd.num_limbs;
  const int b = wd.bits[limb];
  const uint8_t* src = in + p * (long long)wd.limb_off[wd.num_limbs] + wd.limb_off[limb];
  const long long s = j * b;
  uint64_t v;
  if (wd.aligned) {
    const uint64_t* s64 = reinterpret_cast<const uint64_t*>(src);
    const long long w = s >> 6;
    const int off = (int)(s & 63);
    v = s64[w] >> off;
    if (off + b > 64) v |= s64[w + 1] << (64 - off);
  } else {
    const long long k = s >> 3;
    int got = -(int)(s & 7);  // stream position of bit 0 of byte k + t, relative to s
    v = 0;
    for (int t = 0; got < b; ++t, got += 8) {
      const uint64_t byte = src[k + t];
      v |= got < 0 ? byte >> -got : byte << got;
    }
  }
  polys[g] = v & low_mask(b);
}

// ChaCha20 block (RFC 7539 Sec. 2.3) as 16 little-endian 32-bit words.
__device__ __forceinline__ uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
#define SHG_QR(a, b, c, d)                 \
  a += b; d ^= a; d = rotl32(d, 16);       \
  c += d; b ^= c; b = rotl32(b, 12);       \
  a += b; d ^= a; d = rotl32(d, 8);        \
  c += d; b ^= c; b = rotl32(b, 7);
__device__ void chacha_block(const uint32_t (&key)[8], uint32_t counter, uint32_t n0, uint32_t n1,
                             uint32_t n2, uint32_t (&out)[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2],
                    key[3], key[4], key[5], key[6], key[7], counter, n0, n1, n2};
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) x[i] = s[i];
#pragma unroll 1
  for (int r = 0; r < 10; ++r) {
    SHG_QR(x[0], x[4], x[8], x[12]) SHG_QR(x[1], x[5], x[9], x[13])
    SHG_QR(x[2], x[6], x[10], x[14]) SHG_QR(x[3], x[7], x[11], x[15])
    SHG_QR(x[0], x[5], x[10], x[15]) SHG_QR(x[1], x[6], x[11], x[12])
    SHG_QR(x[2], x[7], x[8], x[13]) SHG_QR(x[3], x[4], x[9], x[14])
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}
#undef SHG_QR

constexpr int kPadThreads = 1024;
constexpr long long kWordsPerBuffer = 255 * 32 / 8;  // 1020 draws of 8 bytes per salt
constexpr int kBlocksPerBuffer = 128;                // 127 full blocks and half a block

// Exclusive prefix sum over the workgroup; returns this thread's offset and
// writes the total to *total (both after the call's barriers).
__device__ int block_exclusive_scan(int v, int* wave_sums, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wave_sums[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < kPadThreads / 64; ++i) {
      const int t = wave_sums[i];
      wave_sums[i] = acc;
      acc += t;
    }
    *total = acc;
  }
  __syncthreads();
  return wave_sums[wave] + x - v;
}

// Single workgroup: segments (polynomial p, limb i) in order, each filled
// with n accepted draws.
__global__ void __launch_bounds__(kPadThreads)
    chacha_pad_kernel(PadDesc pd, uint64_t* __restrict__ out) {
  __shared__ int wave_sums[kPadThreads / 64];
  __shared__ int total_acc;
  __shared__ long long next_pos;
  const int t = threadIdx.x;
  const long long n = 1LL << pd.log_n;
  long long wpos = 0;  // index of the next unused 64-bit draw of the stream
  for (long long seg = 0; seg < (long long)pd.count * pd.num_limbs; ++seg) {
    const int limb = (int)(seg % pd.num_limbs);
    const uint64_t q = pd.q[limb];
    const uint64_t m = (1ull << pd.bits[limb]) - 1;
    uint64_t* dst = out + seg * n;
    long long filled = 0;
    while (filled < n) {
      const long long s0 = wpos / kWordsPerBuffer;
      const long long u0 = wpos - s0 * kWordsPerBuffer;
      const long long unit = s0 * kBlocksPerBuffer + u0 / 8 + t;
      const long long salt = unit / kBlocksPerBuffer;
      const int blk = (int)(unit - salt * kBlocksPerBuffer);
      const long long first = salt * kWordsPerBuffer + (long long)blk * 8;
      const int words = blk == kBlocksPerBuffer - 1 ? 4 : 8;
      uint32_t ks[16];
      chacha_block(pd.key, (uint32_t)blk, 0x746c6173u /* "salt" */, (uint32_t)salt,
                   (uint32_t)((unsigned long long)salt >> 32), ks);
      uint64_t val[8];
      unsigned acc_mask = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        val[k] = ((uint64_t)ks[2 * k] | ((uint64_t)ks[2 * k + 1] << 32)) & m;
        const bool ok = k < words && first + k >= wpos && val[k] < q;
        acc_mask |= ok ? (1u << k) : 0u;
      }
      const int mine = __popc(acc_mask);
      const int base = block_exclusive_scan(mine, wave_sums, &total_acc);
      const long long need = n - filled;
      int r = base;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (acc_mask & (1u << k)) {
          if (r < need) dst[filled + r] = val[k];
          if (r == need - 1) next_pos = first + k + 1;  // the segment's last draw
          ++r;
        }
      }
      const int tot = total_acc;
      __syncthreads();  // next_pos written; wave_sums free for the next scan
      if (tot >= need) {
        filled = n;
        wpos = next_pos;
      } else {
        filled += tot;
        const long long end_unit = s0 * kBlocksPerBuffer + u0 / 8 + kPadThreads;
        const long long es = end_unit / kBlocksPerBuffer;
        wpos = es * kWordsPerBuffer + (end_unit - es * kBlocksPerBuffer) * 8;
      }
      __syncthreads();
    }
  }
}

}  // namespace

hipError_t launch_pack(const uint64_t* polys, uint8_t* out, const WireDesc& wd, int count,
                       hipStream_t s) {
  long long max_words = 0;
  for (int i = 0; i < wd.num_limbs; ++i) {
    const long long bytes = (long long)wd.limb_off[i + 1] - (long long)wd.limb_off[i];
    const long long w = (bytes + 7) / 8;
    max_words = w > max_words ? w : max_words;
  }
  const long long total = (long long)count * wd.num_limbs * max_words;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(pack_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, polys, out, wd,
                     max_words, total);
  return hipGetLastError();
}

hipError_t launch_unpack(const uint8_t* in, uint64_t* polys, const WireDesc& wd, int count,
                         hipStream_t s) {
  const long long total = ((long long)count * wd.num_limbs) << wd.log_n;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(unpack_kernel, dim3(grid_of(total, 256)), dim3(256), 0, s, in, polys, wd,
                     total);
  return hipGetLastError();
}

hipError_t launch_chacha_pad(const PadDesc& pd, uint64_t* out, hipStream_t s) {
  if (pd.count == 0) return hipSuccess;
  hipLaunchKernelGGL(chacha_pad_kernel, dim3(1), dim3(kPadThreads), 0, s, pd, out);
  return hipGetLastError();
}

}  // namespace shg
