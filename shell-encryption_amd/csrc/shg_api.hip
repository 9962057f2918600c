// C-ABI of libshellgpu (declared in include/shell_gpu.h): context, device
// memory, prime tables, and the entry points of the RNS hot path.  Argument
// checks mirror the reference's InvalidArgument paths; on error nothing is
// launched and outputs are untouched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "rlwe/montgomery.h"
#include "rlwe/ntt_parameters.h"
#include "shell_gpu.h"
#include "shg_device.h"
#include "shg_kernels.h"

using shg::LimbPrimes;
using shg::PrimeDev;
using ModInt = rlwe::MontgomeryInt<uint64_t>;
using Params = rlwe::MontgomeryIntParams<uint64_t>;

// Work area of shg_relin_apply_noalloc, one per GPU, held from the library's
// load on that GPU until the process ends.  The name is a host-side stand-in
// for the symbol, not a device address: host code uses only the address
// that noalloc_work_address() obtains with hipGetSymbolAddress.
static constexpr size_t kNoAllocWorkBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocWorkBytes];

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(SHG_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

struct PrimeRec {
  uint64_t q = 0;
  int log_n = 0;
  std::unique_ptr<const Params> mp;
  rlwe::NttParameters<ModInt> ntt;
  void* d_tables = nullptr;  // fw, iw (2n u64 each: c, cb pairs), fwd, iwd (n doubles each)
  PrimeDev dev{};
};

}  // namespace

struct shg_context {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<PrimeRec>> primes;
  std::map<std::pair<uint64_t, int>, int> prime_index;
  PrimeDev* d_primes = nullptr;
  int d_capacity = 0;
  void* scratch = nullptr;
  size_t scratch_size = 0;
  // Autotuned NTT configurations: (log_n, limbs per polynomial or -1,
  // total-limb bucket or -1, modulus width class) -> configuration code.
  std::map<std::array<int, 4>, int> tuned;
  // Oblivious expansion / inner product work areas and the NTT-form
  // monomials -X^(n - 2^j) = X^(-2^j), keyed by (prime handles, j).
  struct DevBuf {
    void* p = nullptr;
    size_t size = 0;
  };
  DevBuf exp_s, exp_sub, exp_ptrs, exp_ptrs2, pir_part, ks_planes;
  std::map<std::pair<std::vector<int>, int>, void*> monomials;
  // Captured graphs hold raw pointers to the work buffers above.  While any
  // graph instantiated on this context is alive, a buffer that has to grow
  // is retired here instead of freed, and released once the last graph is
  // destroyed.
  int live_graphs = 0;
  std::vector<void*> retired;
  // Device address of this GPU's no-allocation work area (or null, with
  // the reason in noalloc_error).
  void* noalloc_work = nullptr;
  std::string noalloc_error;
  std::mutex mu;
  // Optional per-kernel timing: HIP events recorded around launches.
  bool profiling = false;
  struct ProfRec {
    std::string name;
    hipEvent_t start, stop;
  };
  std::vector<ProfRec> prof;
  std::map<std::string, std::pair<double, long long>> prof_totals;
};

namespace {

// The address of `scratch` on `device` (which must be current), looked up
// once per device; each GPU holds its own copy of the array.
void* noalloc_work_address(int device, std::string* error) {
  static std::mutex mu;
  static std::map<int, void*> addr;
  static std::map<int, std::string> err;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *error = err[device];
    return it->second;
  }
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    p = nullptr;
    err[device] = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
  }
  addr[device] = p;
  *error = err[device];
  return p;
}

uint64_t mulmod_host(uint64_t a, uint64_t b, uint64_t q) {
  return (uint64_t)((unsigned __int128)a * b % q);
}
uint64_t powmod_host(uint64_t a, uint64_t e, uint64_t q) {
  uint64_t r = 1 % q;
  a %= q;
  for (; e; e >>= 1, a = mulmod_host(a, a, q))
    if (e & 1) r = mulmod_host(r, a, q);
  return r;
}

// The FP64 constants of a ModReduceLsb job (PolyJobDev::f_*) from its
// Montgomery-form ones, for primes q (< 2^50, so exact as doubles):
// f_kl = k * n^-1 * 2^-64 mod q_l, f_c1[i] = r * ql_inv * 2^-128 mod q_i,
// f_c2[i] = t * ql_inv * 2^-64 mod q_i.
void fill_modreduce_f64(shg_context* ctx, shg::PolyJobDev& d, uint64_t k) {
  const int L = d.num_limbs;
  auto rinv = [](uint64_t q) {
    const uint64_t r = (uint64_t)(((unsigned __int128)1 << 64) % q);
    return powmod_host(r, q - 2, q);
  };
  const uint64_t ql = ctx->primes[d.prime[L - 1]]->q;
  const uint64_t ninv = powmod_host((uint64_t)1 << d.log_n, ql - 2, ql);
  d.f_kl = (double)mulmod_host(mulmod_host(k % ql, ninv, ql), rinv(ql), ql);
  for (int i = 0; i < L - 1; ++i) {
    const uint64_t qi = ctx->primes[d.prime[i]]->q;
    const uint64_t ri = rinv(qi);
    const uint64_t base = mulmod_host(d.ql_inv[i], ri, qi);
    d.f_c1[i] = (double)mulmod_host(mulmod_host(d.r_mod_qi[i], base, qi), ri, qi);
    d.f_c2[i] = (double)mulmod_host(d.t_mod_qi[i], base, qi);
  }
}

// True while the context's stream is being captured into a graph.
bool capturing(shg_context* ctx) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(ctx->stream, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

// Entry points that synchronise with the host, allocate or copy from host
// memory cannot be recorded into a graph: they fail before doing anything.
int refuse_capture(shg_context* ctx, const char* what) {
  return fail(SHG_E_FAILED_PRECONDITION, "%s cannot be called while the context's stream is "
              "captured into a graph.", what);
}

// Releases a work buffer that is being replaced: retired while a graph may
// still replay launches that use it, else freed once queued work is done.
int release_work_buffer(shg_context* ctx, void* p) {
  if (!p) return SHG_OK;
  if (ctx->live_graphs > 0) {
    ctx->retired.push_back(p);
    return SHG_OK;
  }
  hipError_t se = hipStreamSynchronize(ctx->stream);  // queued work may still use it
  if (se == hipSuccess) se = hipFree(p);
  return se == hipSuccess ? SHG_OK : hip_fail(se, "releasing a previous work buffer");
}

int grow_refused(size_t bytes) {
  return fail(SHG_E_FAILED_PRECONDITION,
              "this call needs a larger work buffer (%zu bytes) than the context holds, which "
              "cannot be allocated during graph capture; make the same call once outside "
              "capture first.", bytes);
}

int ensure_scratch(shg_context* ctx, size_t bytes) {
  if (bytes <= ctx->scratch_size) return SHG_OK;
  if (capturing(ctx)) return grow_refused(bytes);
  int rc = release_work_buffer(ctx, ctx->scratch);
  if (rc) return rc;
  ctx->scratch = nullptr;
  ctx->scratch_size = 0;
  hipError_t e = hipMalloc(&ctx->scratch, bytes);
  if (e != hipSuccess) return fail(SHG_E_NO_MEMORY, "scratch allocation of %zu bytes failed", bytes);
  ctx->scratch_size = bytes;
  return SHG_OK;
}

int ensure_buf(shg_context* ctx, shg_context::DevBuf& b, size_t bytes) {
  if (bytes <= b.size) return SHG_OK;
  if (capturing(ctx)) return grow_refused(bytes);
  int rc = release_work_buffer(ctx, b.p);
  if (rc) return rc;
  b.p = nullptr;
  b.size = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess)
    return fail(SHG_E_NO_MEMORY, "work buffer allocation of %zu bytes failed", bytes);
  b.size = bytes;
  return SHG_OK;
}

// Records a HIP event pair around the launches made during its lifetime when
// profiling is on.
struct ProfScope {
  shg_context* ctx;
  shg_context::ProfRec rec;
  bool on;
  ProfScope(shg_context* c, const char* name) : ctx(c), on(c && c->profiling && !capturing(c)) {
    if (!on) return;
    rec.name = name;
    if (hipEventCreate(&rec.start) != hipSuccess || hipEventCreate(&rec.stop) != hipSuccess) {
      on = false;
      return;
    }
    (void)hipEventRecord(rec.start, ctx->stream);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.stop, ctx->stream);
    ctx->prof.push_back(rec);
  }
};

void prof_collect(shg_context* ctx) {
  if (ctx->prof.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& r : ctx->prof) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) {
      auto& t = ctx->prof_totals[r.name];
      t.first += ms;
      t.second += 1;
    }
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  ctx->prof.clear();
}

// Validates a prime list and returns log_n and the LimbPrimes table.
int check_primes(shg_context* ctx, const int* primes, int num_limbs, int* log_n, LimbPrimes* lp) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (num_limbs <= 0 || num_limbs > SHG_MAX_LIMBS)
    return fail(SHG_E_INVALID_ARGUMENT, "`moduli` must contain between 1 and %d RNS moduli.",
                SHG_MAX_LIMBS);
  if (!primes) return fail(SHG_E_INVALID_ARGUMENT, "`moduli` must not be empty.");
  int ln = -1;
  for (int i = 0; i < num_limbs; ++i) {
    const int p = primes[i];
    if (p < 0 || p >= (int)ctx->primes.size())
      return fail(SHG_E_INVALID_ARGUMENT, "invalid prime handle %d.", p);
    const int pl = ctx->primes[p]->log_n;
    if (ln < 0) ln = pl;
    if (pl != ln)
      return fail(SHG_E_INVALID_ARGUMENT, "all RNS moduli must share one ring degree.");
    lp->id[i] = p;
  }
  *log_n = ln;
  return SHG_OK;
}

int rows_bucket(long long rows) {
  int bucket = 0;
  while ((1LL << (bucket + 1)) <= rows && bucket < 30) ++bucket;
  return bucket;
}

// NTT launch configuration (shg::launch_ntt) for a call: the explicit
// variant, or for SHG_NTT_AUTO the autotuner's choice for (log_n, rows) if
// there is one, else FP64 when every modulus allows it.  -1: FP64 asked for
// with a modulus of 2^50 or more.
// Width class of a modulus chain for the autotuner's table: 0 when every
// modulus is below 2^50 (the FP64 variants apply), 1 otherwise.
int width_class(shg_context* ctx, const int* primes, int num_limbs) {
  for (int i = 0; i < num_limbs; ++i)
    if (ctx->primes[primes[i]]->q >= (1ull << 50)) return 1;
  return 0;
}

// The autotuned configuration for a shape, or -1: the exact key, then the
// same key for any limbs per polynomial.
int tuned_config(shg_context* ctx, int log_n, int num_limbs, long long rows, int wide) {
  const int bucket = rows_bucket(rows);
  for (const auto& key : {std::array<int, 4>{log_n, num_limbs, bucket, wide},
                          std::array<int, 4>{log_n, -1, bucket, wide}}) {
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) return it->second;
  }
  return -1;
}

int resolve_variant(shg_context* ctx, int variant, const int* primes, int num_limbs, int log_n,
                    long long rows) {
  const int wide = width_class(ctx, primes, num_limbs);
  const int base = variant & 3;
  if (base == SHG_NTT_F64 && wide) return -1;
  if (base == SHG_NTT_INT_SHOUP) return variant & ~SHG_NTT_PREFETCH;
  if (base == SHG_NTT_F64) return variant;
  const int t = tuned_config(ctx, log_n, num_limbs, rows, wide);
  if (t >= 0) return t;
  return wide ? SHG_NTT_INT_SHOUP : SHG_NTT_F64;
}

// Deterministic Miller-Rabin for 64-bit integers (these bases decide every
// n < 3.3 * 10^24).
bool is_prime_u64(uint64_t n) {
  if (n < 2) return false;
  for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull})
    if (n % p == 0) return n == p;
  uint64_t d = n - 1;
  int r = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    ++r;
  }
  auto mulmod = [n](uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % n); };
  for (uint64_t a : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
    uint64_t x = 1, base = a % n, e = d;
    while (e) {
      if (e & 1) x = mulmod(x, base);
      base = mulmod(base, base);
      e >>= 1;
    }
    if (x == 1 || x == n - 1) continue;
    bool composite = true;
    for (int i = 1; i < r; ++i) {
      x = mulmod(x, x);
      if (x == n - 1) {
        composite = false;
        break;
      }
    }
    if (composite) return false;
  }
  return true;
}

bool aligned_ptr(const void* p) { return p != nullptr && ((uintptr_t)p % 8) == 0; }
// Transform and key-switch kernels move residues in 16-byte pairs.
bool aligned16(const void* p) { return p != nullptr && ((uintptr_t)p % 16) == 0; }

int build_gadget(shg_context* ctx, const int* primes, int num_limbs, const int* log_bs,
                 shg::GadgetDesc* gd) {
  if (!log_bs) return fail(SHG_E_INVALID_ARGUMENT, "`log_bs` must contain %d elements.", num_limbs);
  int k = 0, max_b = 0;
  uint64_t min_q = ~0ull;
  for (int i = 0; i < num_limbs; ++i) {
    const int b = log_bs[i];
    const unsigned lq = ctx->primes[primes[i]]->mp->log_modulus;
    if (b <= 0 || b >= 64) return fail(SHG_E_INVALID_ARGUMENT, "gadget base log must be in [1, 63].");
    const int dims = (int)((lq + (b - 1)) / b);  // GadgetSize (gadget.h:43-47)
    for (int d = 0; d < dims; ++d) {
      if (k >= 256) return fail(SHG_E_UNSUPPORTED, "gadget dimension above 256.");
      gd->src_limb[k] = (unsigned char)i;
      gd->shift[k] = (unsigned char)(d * b);
      gd->log_b[k] = (unsigned char)b;
      ++k;
    }
    gd->limb_first[i] = (unsigned char)(k - dims);
    gd->limb_digits[i] = (unsigned char)dims;
    gd->limb_log_b[i] = (unsigned char)b;
    max_b = b > max_b ? b : max_b;
    const uint64_t q = ctx->primes[primes[i]]->q;
    // Largest digit of this limb: min(2^b - 1, q - 1).
    const uint64_t max_digit = ((1ull << b) - 1) < q - 1 ? ((1ull << b) - 1) : q - 1;
    gd->limb_balance[i] = max_digit > (q >> 1) ? 1 : 0;
    gd->any_balance |= gd->limb_balance[i];
    min_q = q < min_q ? q : min_q;
  }
  gd->k = k;
  gd->max_log_b = max_b;
  bool uniform = num_limbs > 0;
  for (int i = 1; i < num_limbs; ++i) uniform &= log_bs[i] == log_bs[0];
  gd->plane_bits = uniform && (max_b == 8 || max_b == 10 || max_b == 12) ? max_b : 0;
  gd->small_digits = max_b + 6 < 64 && (1ull << (max_b + 6)) <= min_q;
  return SHG_OK;
}

}  // namespace

namespace shg {
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace shg

extern "C" {

const char* shg_last_error(void) { return g_last_error.c_str(); }

int shg_create(int device, shg_context** out) {
  if (!out) return fail(SHG_E_INVALID_ARGUMENT, "out must not be null.");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(SHG_E_DEVICE, "no HIP device available.");
  if (device < 0 || device >= n) return fail(SHG_E_INVALID_ARGUMENT, "device %d out of range.", device);
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  auto* ctx = new shg_context();
  ctx->device = device;
  e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return hip_fail(e, "hipStreamCreate");
  }
  ctx->stream = ctx->own_stream;
  ctx->noalloc_work = noalloc_work_address(device, &ctx->noalloc_error);
  ctx->d_capacity = 256;
  e = hipMalloc(&ctx->d_primes, sizeof(PrimeDev) * ctx->d_capacity);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(ctx->own_stream);  // already failing; report the allocation
    delete ctx;
    return hip_fail(e, "hipMalloc(prime table)");
  }
  *out = ctx;
  return SHG_OK;
}

int shg_destroy(shg_context* ctx) {
  if (!ctx) return SHG_OK;
  // Teardown releases everything it can; the first failure is reported.
  hipError_t first = hipSuccess;
  auto note = [&](hipError_t e) {
    if (first == hipSuccess) first = e;
  };
  note(hipSetDevice(ctx->device));
  note(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->primes)
    if (p->d_tables) note(hipFree(p->d_tables));
  if (ctx->d_primes) note(hipFree(ctx->d_primes));
  if (ctx->scratch) note(hipFree(ctx->scratch));
  for (auto* b : {&ctx->exp_s, &ctx->exp_sub, &ctx->exp_ptrs, &ctx->exp_ptrs2, &ctx->pir_part,
                  &ctx->ks_planes})
    if (b->p) note(hipFree(b->p));
  for (auto& m : ctx->monomials) note(hipFree(m.second));
  for (void* p : ctx->retired) note(hipFree(p));
  if (ctx->own_stream) note(hipStreamDestroy(ctx->own_stream));
  delete ctx;
  return first == hipSuccess ? SHG_OK : hip_fail(first, "context teardown");
}

int shg_set_stream(shg_context* ctx, void* stream) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
  return SHG_OK;
}

void* shg_get_stream(shg_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int shg_sync(shg_context* ctx) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_sync");
  hipError_t e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "hipStreamSynchronize");
}

int shg_alloc(shg_context* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_alloc");
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 8);
  return e == hipSuccess ? SHG_OK : fail(SHG_E_NO_MEMORY, "hipMalloc(%zu) failed", bytes);
}

int shg_free(shg_context* ctx, void* dptr) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (!dptr) return SHG_OK;
  if (capturing(ctx)) return refuse_capture(ctx, "shg_free");
  // Work queued on the stream may still use the buffer.
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipFree(dptr);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "hipFree");
}

int shg_upload(shg_context* ctx, void* dptr, const void* host, size_t bytes) {
  if (!ctx || (!dptr && bytes) || (!host && bytes)) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_upload");
  hipError_t e = hipMemcpyAsync(dptr, host, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "upload");
}

int shg_download(shg_context* ctx, void* host, const void* dptr, size_t bytes) {
  if (!ctx || (!dptr && bytes) || (!host && bytes)) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_download");
  hipError_t e = hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "download");
}

int shg_copy(shg_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (!bytes || dst == src) return SHG_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "copy");
}

int shg_prime_create(shg_context* ctx, uint64_t q, int log_n, int* prime) {
  if (!ctx || !prime) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_prime_create");
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto key = std::make_pair(q, log_n);
  auto it = ctx->prime_index.find(key);
  if (it != ctx->prime_index.end()) {
    *prime = it->second;
    return SHG_OK;
  }
  auto mp_or = Params::Create(q);
  if (!mp_or.ok()) return fail(SHG_E_INVALID_ARGUMENT, "%s", mp_or.status().message().c_str());
  auto rec = std::make_unique<PrimeRec>();
  rec->q = q;
  rec->log_n = log_n;
  rec->mp = std::move(mp_or).value();
  auto ntt_or = rlwe::InitializeNttParameters<ModInt>(log_n, rec->mp.get());
  if (!ntt_or.ok()) return fail(SHG_E_INVALID_ARGUMENT, "%s", ntt_or.status().message().c_str());
  rec->ntt = std::move(ntt_or).value();
  const size_t n = (size_t)1 << log_n;
  std::vector<uint64_t> fw(2 * n), iw(2 * n);
  std::vector<double> fwd(n), iwd(n);
  const double qd = (double)q;
  for (size_t i = 0; i < n; ++i) {
    fw[2 * i] = std::get<0>(rec->ntt.psis_bitrev_constant[i]);
    fw[2 * i + 1] = std::get<1>(rec->ntt.psis_bitrev_constant[i]);
    iw[2 * i] = std::get<0>(rec->ntt.psis_inv_bitrev_constant[i]);
    iw[2 * i + 1] = std::get<1>(rec->ntt.psis_inv_bitrev_constant[i]);
    // FP64 twiddles in balanced form (w - q for w > q/2): |w| <= q/2
    // halves the quotient error of every FP64 product, which the reduction
    // schedules rely on (shg_ntt_core.h, shg_keyswitch.hip).
    fwd[i] = fw[2 * i] > q / 2 ? (double)fw[2 * i] - qd : (double)fw[2 * i];
    iwd[i] = iw[2 * i] > q / 2 ? (double)iw[2 * i] - qd : (double)iw[2 * i];
  }
  const size_t bytes = 6 * n * 8;
  hipError_t e = hipMalloc(&rec->d_tables, bytes);
  if (e != hipSuccess) return fail(SHG_E_NO_MEMORY, "twiddle table allocation failed");
  char* base = (char*)rec->d_tables;
  e = hipMemcpy(base, fw.data(), 2 * n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(base + 2 * n * 8, iw.data(), 2 * n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(base + 4 * n * 8, fwd.data(), n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(base + 5 * n * 8, iwd.data(), n * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "twiddle upload");
  PrimeDev& d = rec->dev;
  const Params& P = *rec->mp;
  d.q = q;
  d.inv_modulus = P.inv_modulus;
  d.r_mod_q = P.r_mod_modulus;
  d.r_barrett = P.r_mod_modulus_barrett;
  d.inv_r = P.inv_r;
  d.inv_r_barrett = P.inv_r_barrett;
  d.bn = (uint64_t)P.barrett_numerator;
  auto ninv_c = rec->ntt.n_inv_ptr->GetConstant(&P);
  d.n_inv_c = std::get<0>(ninv_c);
  d.n_inv_cb = std::get<1>(ninv_c);
  d.qd = qd;
  d.qinv = 1.0 / qd;
  d.n_inv_d = (double)d.n_inv_c;
  d.n_inv_dq = (double)d.n_inv_c / qd;
  d.n_inv_x_c = P.ExportInt(d.n_inv_c);  // n^-1 * 2^-64 mod q
  d.n_inv_x_cb = (uint64_t)(((rlwe::Uint128)d.n_inv_x_c << 64) / q);
  d.n_inv_x_d = (double)d.n_inv_x_c;
  d.n_inv_x_dq = (double)d.n_inv_x_c / qd;
  d.log_n = log_n;
  d.log_q = (int)P.log_modulus;
  d.fw = (const uint64_t*)base;
  d.iw = (const uint64_t*)(base + 2 * n * 8);
  d.fwd = (const double*)(base + 4 * n * 8);
  d.iwd = (const double*)(base + 5 * n * 8);
  const int id = (int)ctx->primes.size();
  if (id >= ctx->d_capacity) {
    PrimeDev* bigger = nullptr;
    e = hipMalloc(&bigger, sizeof(PrimeDev) * ctx->d_capacity * 2);
    if (e != hipSuccess) return fail(SHG_E_NO_MEMORY, "prime table growth failed");
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpy(bigger, ctx->d_primes, sizeof(PrimeDev) * ctx->d_capacity,
                    hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipFree(ctx->d_primes);
    if (e != hipSuccess) {
      (void)hipFree(bigger);  // already failing; report the copy
      return hip_fail(e, "prime table growth");
    }
    ctx->d_primes = bigger;
    ctx->d_capacity *= 2;
  }
  e = hipMemcpy(ctx->d_primes + id, &d, sizeof(PrimeDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "prime upload");
  ctx->primes.push_back(std::move(rec));
  ctx->prime_index[key] = id;
  *prime = id;
  return SHG_OK;
}

int shg_rns_ql_inv(shg_context* ctx, const int* primes, int num_limbs, uint64_t* host_ql_inv) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (num_limbs < 2)
    return fail(SHG_E_INVALID_ARGUMENT, "ModReduceLsb cannot apply with only one prime modulus.");
  if (!host_ql_inv) return fail(SHG_E_INVALID_ARGUMENT, "`ql_inv` must not be null.");
  // q_l^-1 mod q_i (RnsContext::GenerateCrtConstantsForMainModulus,
  // rns/rns_context.cc:169-199): ImportInt(q_l) then its inverse.
  const uint64_t ql = ctx->primes[primes[num_limbs - 1]]->q;
  for (int i = 0; i < num_limbs - 1; ++i) {
    const Params* Pi = ctx->primes[primes[i]]->mp.get();
    if (Pi->modulus == ql)
      return fail(SHG_E_INVALID_ARGUMENT, "the RNS moduli must be distinct.");
    auto v = ModInt::ImportInt(ql, Pi);
    if (!v.ok()) return fail(SHG_E_INVALID_ARGUMENT, "%s", v.status().message().c_str());
    host_ql_inv[i] = v.value().MultiplicativeInverse(Pi).GetMontgomeryRepresentation();
  }
  return SHG_OK;
}

int shg_prime_info(shg_context* ctx, int prime, uint64_t* params, uint64_t* tables) {
  if (!ctx || prime < 0 || prime >= (int)ctx->primes.size())
    return fail(SHG_E_INVALID_ARGUMENT, "invalid prime handle.");
  const PrimeRec& r = *ctx->primes[prime];
  const Params& P = *r.mp;
  if (params) {
    params[0] = P.modulus;
    params[1] = P.log_modulus;
    params[2] = P.r_mod_modulus;
    params[3] = P.r_mod_modulus_barrett;
    params[4] = P.inv_modulus;
    params[5] = P.inv_r;
    params[6] = P.inv_r_barrett;
    params[7] = (uint64_t)P.barrett_numerator;
    params[8] = r.ntt.psi->GetMontgomeryRepresentation();
    params[9] = r.ntt.n_inv_ptr->GetMontgomeryRepresentation();
  }
  if (tables) {
    const size_t n = (size_t)1 << r.log_n;
    for (size_t i = 0; i < n; ++i) {
      tables[i] = std::get<0>(r.ntt.psis_bitrev_constant[i]);
      tables[n + i] = std::get<1>(r.ntt.psis_bitrev_constant[i]);
      tables[2 * n + i] = std::get<0>(r.ntt.psis_inv_bitrev_constant[i]);
      tables[3 * n + i] = std::get<1>(r.ntt.psis_inv_bitrev_constant[i]);
    }
  }
  return SHG_OK;
}

static int ntt_common(shg_context* ctx, bool inverse, uint64_t* polys, int count, int num_limbs,
                      const int* primes, int variant) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (count < 0) return fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (count == 0) return SHG_OK;
  if (!aligned16(polys))
    return fail(SHG_E_INVALID_ARGUMENT, "polynomial data must be a non-null 16-byte aligned pointer.");
  if (variant < 0 || (variant & 3) == 3 || (variant & ~255) != 0)
    return fail(SHG_E_INVALID_ARGUMENT, "unknown transform variant %d.", variant);
  const long long rows = (long long)count * num_limbs;
  if ((rows << (log_n > shg::kMaxLogM ? log_n - shg::kMaxLogM : 0)) > 0x7fffffffLL)
    return fail(SHG_E_INVALID_ARGUMENT, "batch too large for one launch.");
  const int v = resolve_variant(ctx, variant, primes, num_limbs, log_n, rows);
  if (v < 0) return fail(SHG_E_INVALID_ARGUMENT, "FP64 transform requires moduli below 2^50.");
  ProfScope prof(ctx, inverse ? "ntt_inv" : "ntt_fwd");
  hipError_t e = shg::launch_ntt(inverse, v, polys, ctx->d_primes, lp, (int)rows, num_limbs, log_n,
                                 shg::kMaxLogM, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "ntt launch");
}

int shg_ntt_fwd(shg_context* ctx, uint64_t* polys, int count, int num_limbs, const int* primes,
                int variant) {
  return ntt_common(ctx, false, polys, count, num_limbs, primes, variant);
}

int shg_ntt_inv(shg_context* ctx, uint64_t* polys, int count, int num_limbs, const int* primes,
                int variant) {
  return ntt_common(ctx, true, polys, count, num_limbs, primes, variant);
}

int shg_pointwise(shg_context* ctx, int op, uint64_t* x, const uint64_t* a, size_t a_stride,
                  const uint64_t* b, size_t b_stride, int count, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (op < SHG_OP_ADD || op > SHG_OP_FMA) return fail(SHG_E_INVALID_ARGUMENT, "unknown op %d.", op);
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(x)) return fail(SHG_E_INVALID_ARGUMENT, "`this` must not be null.");
  if (op != SHG_OP_NEG && !aligned_ptr(a)) return fail(SHG_E_INVALID_ARGUMENT, "`that` must not be null.");
  if (op == SHG_OP_FMA && !aligned_ptr(b)) return fail(SHG_E_INVALID_ARGUMENT, "`b` must not be null.");
  ProfScope prof(ctx, "pointwise");
  hipError_t e = shg::launch_pointwise(op, x, a, a_stride, b, b_stride, ctx->d_primes, lp, count,
                                       num_limbs, log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "pointwise launch");
}

int shg_fused_mul_sum_add(shg_context* ctx, int subtract, uint64_t* x, const uint64_t* a,
                          const uint64_t* b, const uint64_t* c, int count, int num_limbs,
                          const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(x)) return fail(SHG_E_INVALID_ARGUMENT, "`this` must not be null.");
  if (!aligned_ptr(a) || !aligned_ptr(b) || !aligned_ptr(c))
    return fail(SHG_E_INVALID_ARGUMENT, "Polynomials `a`, `b`, and `c` must not be null.");
  // rns/lazy_rns_polynomial.cc:73-78: a + b (or a + q - b) and its product
  // with c must fit the reference's 128-bit lazy sum.
  for (int i = 0; i < num_limbs; ++i)
    if (ctx->primes[primes[i]]->q >> 60)
      return fail(SHG_E_INVALID_ARGUMENT, "Modulus is too large to perform fused multiply-sum-add.");
  ProfScope prof(ctx, "pointwise");
  hipError_t e = shg::launch_fused_mul_sum_add(subtract != 0, x, a, b, c, ctx->d_primes, lp, count,
                                               num_limbs, log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "fused multiply-sum-add launch");
}

int shg_mul_scalar(shg_context* ctx, uint64_t* x, const uint64_t* host_scalars, int count,
                   int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (!host_scalars)
    return fail(SHG_E_INVALID_ARGUMENT, "`scalar_mod_qs` must contain %d modular integers.", num_limbs);
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(x)) return fail(SHG_E_INVALID_ARGUMENT, "`this` must not be null.");
  LimbPrimes lo{}, hi{};
  for (int i = 0; i < num_limbs; ++i) {
    lo.id[i] = (int)(uint32_t)host_scalars[i];
    hi.id[i] = (int)(uint32_t)(host_scalars[i] >> 32);
  }
  hipError_t e = shg::launch_mul_scalar(x, lo, hi, ctx->d_primes, lp, count, num_limbs, log_n,
                                        ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "mul_scalar launch");
}

int shg_mul_int(shg_context* ctx, uint64_t* x, uint64_t scalar, int count, int num_limbs,
                const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  std::vector<uint64_t> s(num_limbs);
  for (int i = 0; i < num_limbs; ++i)
    s[i] = ModInt::ImportInt(scalar, ctx->primes[primes[i]]->mp.get()).value().GetMontgomeryRepresentation();
  return shg_mul_scalar(ctx, x, s.data(), count, num_limbs, primes);
}

int shg_substitute(shg_context* ctx, const uint64_t* in, uint64_t* out, int power, int count,
                   int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  const int n = 1 << log_n;
  if (power < 0 || (power % 2) == 0 || power >= 2 * n)
    return fail(SHG_E_INVALID_ARGUMENT,
                "Substitution power must be a non-negative odd integer less than 2*n.");
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(in) || !aligned_ptr(out) || in == out)
    return fail(SHG_E_INVALID_ARGUMENT, "substitute needs distinct input and output buffers.");
  ProfScope prof(ctx, "substitute");
  hipError_t e = shg::launch_substitute(in, out, power, count, num_limbs, log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "substitute launch");
}

}  // extern "C"

// ModReduceLsb and ModReduceMsb (rns/rns_polynomial.cc:262-417) share one
// form: limb i <- (r * a_i + s * lift([k * a_l])) * q_l^-1 mod q_i, with
// Lsb: k = (q_l - r) / t, r = q_l mod t, s = t;  Msb: k = 1, r = 1, s = -1.
// r_mod_qi and s_mod_qi are Montgomery reps per kept limb.
static int mod_reduce_impl(shg_context* ctx, uint64_t* polys, uint64_t k,
                           const uint64_t* r_mod_qi, const uint64_t* s_mod_qi,
                           const uint64_t* host_ql_inv, int is_ntt, int count, int num_limbs,
                           const int* primes, const LimbPrimes& lp, int log_n) {
  int rc = SHG_OK;
  const int nr = num_limbs - 1;
  const int last = primes[nr];
  const Params* Pl = ctx->primes[last]->mp.get();
  shg::ModReduceArgs args{};
  args.k_mod_ql = ModInt::ImportInt(k, Pl).value().GetMontgomeryRepresentation();
  for (int i = 0; i < nr; ++i) {
    args.r_mod_qi[i] = r_mod_qi[i];
    args.t_mod_qi[i] = s_mod_qi[i];
    args.ql_inv[i] = host_ql_inv[i];
  }
  const size_t n = (size_t)1 << log_n;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (log_n >= shg::kBatchMinLogN && log_n <= shg::kBatchMaxLogN &&
      num_limbs <= shg::kBatchMaxPolyLimbs) {
    // One workgroup per polynomial: INTT of the last limb into its scratch,
    // then lift, NTT and combine for each other limb, all in one launch.
    const int v = resolve_variant(ctx, SHG_NTT_AUTO, primes, num_limbs, log_n,
                                  (long long)count * num_limbs) & 3;
    shg::PolyJobDev uni{};
    uni.data = polys;
    uni.log_n = log_n;
    uni.num_limbs = num_limbs;
    uni.is_ntt = is_ntt ? 1 : 0;
    for (int i = 0; i < num_limbs; ++i) uni.prime[i] = primes[i];
    uni.k_mod_ql = args.k_mod_ql;
    for (int i = 0; i < nr; ++i) {
      uni.r_mod_qi[i] = args.r_mod_qi[i];
      uni.t_mod_qi[i] = args.t_mod_qi[i];
      uni.ql_inv[i] = args.ql_inv[i];
    }
    if (v == SHG_NTT_F64) fill_modreduce_f64(ctx, uni, k);
    if (v == SHG_NTT_F64 && log_n == 14 && is_ntt && count * (long long)num_limbs <= 0x7fffffffLL) {
      // Two launches on the two-workgroups-per-CU transforms.
      rc = ensure_scratch(ctx, ((size_t)count << 14) * 8);
      if (rc) return rc;
      hipError_t e = shg::launch_modreduce_w32(uni, (size_t)num_limbs * n, count, ctx->d_primes,
                                               (uint64_t*)ctx->scratch, ctx->stream);
      return e == hipSuccess ? SHG_OK : hip_fail(e, "mod_reduce launch");
    }
    const int grid = shg::modreduce_uniform_grid(v, log_n, count);
    rc = ensure_scratch(ctx, ((size_t)grid << 16) * 8);
    if (rc) return rc;
    hipError_t e = shg::launch_modreduce_uniform(uni, (size_t)num_limbs * n, count, v,
                                                 ctx->d_primes, (uint64_t*)ctx->scratch, grid,
                                                 ctx->stream);
    return e == hipSuccess ? SHG_OK : hip_fail(e, "mod_reduce launch");
  }
  const size_t need = (size_t)count * n * 8 * (1 + nr);
  rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  uint64_t* delta_l = (uint64_t*)ctx->scratch;
  uint64_t* delta = delta_l + (size_t)count * n;
  hipStream_t s = ctx->stream;
  hipError_t e = shg::launch_modreduce_extract(polys, delta_l, ctx->d_primes, last, args.k_mod_ql,
                                               count, num_limbs, log_n, s);
  LimbPrimes lpl{};
  lpl.id[0] = last;
  LimbPrimes lpr = lp;
  if (e == hipSuccess && is_ntt)
    e = shg::launch_ntt(true, resolve_variant(ctx, SHG_NTT_AUTO, &last, 1, log_n, count), delta_l,
                        ctx->d_primes, lpl, count, 1, log_n, shg::kMaxLogM, s);
  if (e == hipSuccess)
    e = shg::launch_modreduce_lift(delta_l, delta, ctx->d_primes, last, lpr, count, num_limbs,
                                   log_n, s);
  if (e == hipSuccess && is_ntt)
    e = shg::launch_ntt(false, resolve_variant(ctx, SHG_NTT_AUTO, primes, nr, log_n,
                                               (long long)count * nr),
                        delta, ctx->d_primes, lpr, count * nr, nr, log_n, shg::kMaxLogM, s);
  if (e == hipSuccess)
    e = shg::launch_modreduce_combine(polys, delta, args, ctx->d_primes, lpr, count, num_limbs,
                                      log_n, s);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "mod_reduce launch");
}

static int mod_reduce_checks(shg_context* ctx, uint64_t* polys, const uint64_t* host_ql_inv,
                             int count, int num_limbs, const int* primes, const char* name,
                             LimbPrimes* lp, int* log_n, bool* done) {
  *done = true;
  int rc = check_primes(ctx, primes, num_limbs, log_n, lp);
  if (rc) return rc;
  if (num_limbs <= 1)
    return fail(SHG_E_INVALID_ARGUMENT, "%s cannot apply with only one prime modulus.", name);
  if (!host_ql_inv)
    return fail(SHG_E_INVALID_ARGUMENT, "`ql_inv` must be defined with respect to %d moduli.",
                num_limbs - 1);
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(polys)) return fail(SHG_E_INVALID_ARGUMENT, "polynomial data must not be null.");
  *done = false;
  return SHG_OK;
}

extern "C" {

int shg_mod_reduce_lsb(shg_context* ctx, uint64_t* polys, uint64_t t, const uint64_t* host_ql_inv,
                       int is_ntt, int count, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  bool done = true;
  if (t == 0 && ctx) return fail(SHG_E_INVALID_ARGUMENT, "plaintext modulus must be positive.");
  int rc = mod_reduce_checks(ctx, polys, host_ql_inv, count, num_limbs, primes, "ModReduceLsb",
                             &lp, &log_n, &done);
  if (rc || done) return rc;
  // rns/rns_polynomial.cc:286-300: q_l = k * t + r.
  const int nr = num_limbs - 1;
  const uint64_t ql = ctx->primes[primes[nr]]->q;
  const uint64_t r = ql % t;
  const uint64_t k = (ql - r) / t;
  uint64_t rq[SHG_MAX_LIMBS], tq[SHG_MAX_LIMBS];
  for (int i = 0; i < nr; ++i) {
    const Params* Pi = ctx->primes[primes[i]]->mp.get();
    rq[i] = ModInt::ImportInt(r, Pi).value().GetMontgomeryRepresentation();
    tq[i] = ModInt::ImportInt(t, Pi).value().GetMontgomeryRepresentation();
  }
  return mod_reduce_impl(ctx, polys, k, rq, tq, host_ql_inv, is_ntt, count, num_limbs, primes,
                         lp, log_n);
}

int shg_mod_reduce_msb(shg_context* ctx, uint64_t* polys, const uint64_t* host_ql_inv, int is_ntt,
                       int count, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  bool done = true;
  int rc = mod_reduce_checks(ctx, polys, host_ql_inv, count, num_limbs, primes, "ModReduceMsb",
                             &lp, &log_n, &done);
  if (rc || done) return rc;
  // (a - lift([a]_{q_l})) * q_l^-1: k = 1, r = 1, s = -1.
  const int nr = num_limbs - 1;
  uint64_t one[SHG_MAX_LIMBS], minus_one[SHG_MAX_LIMBS];
  for (int i = 0; i < nr; ++i) {
    const Params* Pi = ctx->primes[primes[i]]->mp.get();
    one[i] = Pi->r_mod_modulus;
    minus_one[i] = Pi->modulus - Pi->r_mod_modulus;
  }
  return mod_reduce_impl(ctx, polys, 1, one, minus_one, host_ql_inv, is_ntt, count, num_limbs,
                         primes, lp, log_n);
}

int shg_switch_rns_basis(shg_context* ctx, int mode, const uint64_t* in, int in_is_ntt,
                         int balanced, int count, int num_in, const int* in_primes, uint64_t* out,
                         int num_out, const int* out_primes, const uint64_t* host_q_hat_inv_mod_qs,
                         const uint64_t* host_matrix, const uint64_t* host_vec) {
  LimbPrimes lin{}, lout{};
  int log_n = 0, log_n_out = 0;
  int rc = check_primes(ctx, in_primes, num_in, &log_n, &lin);
  if (rc) return rc;
  rc = check_primes(ctx, out_primes, num_out, &log_n_out, &lout);
  if (rc) return rc;
  if (log_n != log_n_out)
    return fail(SHG_E_INVALID_ARGUMENT, "both RNS chains must share one ring degree.");
  if (mode != SHG_BASIS_EXACT && mode != SHG_BASIS_SCALE && mode != SHG_BASIS_APPROX)
    return fail(SHG_E_INVALID_ARGUMENT, "unknown basis conversion %d.", mode);
  if (num_in > shg::kBasisMax || num_out > shg::kBasisMax)
    return fail(SHG_E_UNSUPPORTED, "basis conversions take at most %d moduli per chain.",
                shg::kBasisMax);
  if (in_is_ntt && mode != SHG_BASIS_APPROX)
    return fail(SHG_E_INVALID_ARGUMENT, "Polynomial must be in coefficient form.");
  if (!host_q_hat_inv_mod_qs || !host_matrix || (mode != SHG_BASIS_APPROX && !host_vec))
    return fail(SHG_E_INVALID_ARGUMENT, "the CRT constants must not be null.");
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(in) || !aligned_ptr(out))
    return fail(SHG_E_INVALID_ARGUMENT, "polynomial data must not be null.");
  const size_t n = (size_t)1 << log_n;
  if ((const void*)out < (const void*)(in + (size_t)count * num_in * n) &&
      (const void*)in < (const void*)(out + (size_t)count * num_out * n))
    return fail(SHG_E_INVALID_ARGUMENT, "`out` must not overlap the input.");
  shg::BasisArgs ba{};
  ba.mode = mode;
  ba.nin = num_in;
  ba.nout = num_out;
  ba.balanced = mode == SHG_BASIS_APPROX && balanced ? 1 : 0;
  for (int i = 0; i < num_in; ++i) {
    ba.in_id[i] = in_primes[i];
    ba.q_hat_inv[i] = host_q_hat_inv_mod_qs[i];
    if (mode == SHG_BASIS_SCALE) ba.pre[i] = host_vec[i];
  }
  for (int j = 0; j < num_out; ++j) {
    ba.out_id[j] = out_primes[j];
    if (mode == SHG_BASIS_EXACT) ba.vec[j] = host_vec[j];
  }
  for (int i = 0; i < num_in; ++i)
    for (int j = 0; j < num_out; ++j) {
      if (mode == SHG_BASIS_SCALE) {
        // -(q_i^-1 mod p_j): NegateInPlace of the reference (montgomery.h:390-399).
        const uint64_t v = host_matrix[(size_t)j * num_in + i];
        ba.mat[i * num_out + j] = ctx->primes[out_primes[j]]->q - v;
      } else {
        ba.mat[i * num_out + j] = host_matrix[(size_t)i * num_out + j];
      }
    }
  std::lock_guard<std::mutex> lock(ctx->mu);
  const uint64_t* src = in;
  hipError_t e = hipSuccess;
  if (in_is_ntt) {
    // The inverse transform commutes with the per-limb scaling by
    // (Q/q_i)^-1, so the kernel reads coefficient-form input.
    rc = ensure_scratch(ctx, (size_t)count * num_in * n * 8);
    if (rc) return rc;
    e = hipMemcpyAsync(ctx->scratch, in, (size_t)count * num_in * n * 8, hipMemcpyDeviceToDevice,
                       ctx->stream);
    if (e == hipSuccess)
      e = shg::launch_ntt(true, resolve_variant(ctx, SHG_NTT_AUTO, in_primes, num_in, log_n,
                                                (long long)count * num_in),
                          (uint64_t*)ctx->scratch, ctx->d_primes, lin, count * num_in, num_in,
                          log_n, shg::kMaxLogM, ctx->stream);
    src = (const uint64_t*)ctx->scratch;
  }
  if (e == hipSuccess)
    e = shg::launch_basis_switch(src, out, ba, ctx->d_primes, count, log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "basis conversion launch");
}

int shg_tensor_product(shg_context* ctx, const uint64_t* a, int len_a, const uint64_t* b, int len_b,
                       uint64_t* out, int count, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (len_a <= 0 || len_b <= 0 || len_a > 8 || len_b > 8)
    return fail(SHG_E_INVALID_ARGUMENT, "ciphertext lengths must be in [1, 8].");
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(a) || !aligned_ptr(b) || !aligned_ptr(out))
    return fail(SHG_E_INVALID_ARGUMENT, "ciphertext data must not be null.");
  ProfScope prof(ctx, "tensor_product");
  hipError_t e = shg::launch_tensor_product(a, len_a, b, len_b, out, ctx->d_primes, lp, count,
                                            num_limbs, log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "tensor launch");
}

int shg_gadget_decompose(shg_context* ctx, const uint64_t* a, const int* host_log_bs,
                         uint64_t* digits, int count, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  shg::GadgetDesc gd{};
  rc = build_gadget(ctx, primes, num_limbs, host_log_bs, &gd);
  if (rc) return rc;
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned_ptr(a) || !aligned_ptr(digits))
    return fail(SHG_E_INVALID_ARGUMENT, "data must not be null.");
  hipError_t e = shg::launch_gadget_decompose(a, digits, gd, ctx->d_primes, lp, count, num_limbs,
                                              log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "decompose launch");
}

// One source component of a fused key switch (4 <= log n <= 14) for `count`
// ciphertexts: its gadget digits as 16-bit planes in `planes` (NTT-form
// source, when `planes` is given), or an exported coefficient-form copy in
// `src`, then the key-switch kernel, which initialises out_c0/out_c1 from
// c0/c1 or accumulates into them.
static hipError_t keyswitch_fused_one(shg_context* ctx, int v, uint16_t* planes, uint64_t* src,
                                      const uint64_t* comps, int comps_per_ct, int comp,
                                      bool comp_is_ntt, int key_base, bool init_c1,
                                      bool accumulate, const uint64_t* key_as,
                                      const uint64_t* key_bs, uint64_t* out_c0, uint64_t* out_c1,
                                      int count, int num_limbs, const LimbPrimes& lp, int log_n,
                                      const shg::GadgetDesc& gd,
                                      const uint64_t* f64_as = nullptr,
                                      const uint64_t* f64_bs = nullptr) {
  const size_t poly_words = (size_t)num_limbs << log_n;
  const size_t ct_words = (size_t)comps_per_ct * poly_words;
  hipStream_t s = ctx->stream;
  hipError_t e = hipSuccess;
  shg::KeySwitchArgs ka{};
  if (planes) {
    ProfScope prof(ctx, "keyswitch_intt");
    e = shg::launch_ntt_inv_digits(v, comps + (size_t)comp * poly_words, ct_words, planes, gd,
                                   ctx->d_primes, lp, count, num_limbs, log_n, s);
    ka.planes = planes;
  } else {
    e = hipMemcpy2DAsync(src, poly_words * 8, comps + (size_t)comp * poly_words, ct_words * 8,
                         poly_words * 8, count, hipMemcpyDeviceToDevice, s);
    // The kernel reads exported residues: ExportInt folded into the inverse
    // transform's scaling, or applied in place to coefficient-form input.
    if (e == hipSuccess && comp_is_ntt) {
      ProfScope prof(ctx, "keyswitch_intt");
      e = shg::launch_ntt(true, v, src, ctx->d_primes, lp, count * num_limbs, num_limbs, log_n,
                          shg::kMaxLogM, s, 2);
    } else if (e == hipSuccess) {
      e = shg::launch_export(src, ctx->d_primes, lp, count, num_limbs, log_n, s);
    }
    ka.src = src;
    ka.src_stride = poly_words;
  }
  if (e != hipSuccess) return e;
  ka.c0 = comps;
  ka.c0_stride = ct_words;
  ka.c1 = init_c1 ? comps + poly_words : nullptr;
  ka.c1_stride = ct_words;
  // FP64-form copies of the keys (shg_ks_key) serve the digit-plane kernels
  // at 2^12..2^14; every other path reads the Montgomery words.
  const bool f64_keys = f64_as && f64_bs && planes && (v & 3) == SHG_NTT_F64 && gd.small_digits &&
                        log_n >= 12;
  ka.key_a = f64_keys ? f64_as : key_as;
  ka.key_b = f64_keys ? f64_bs : key_bs;
  ka.key_f64 = f64_keys ? 1 : 0;
  ka.key_base = key_base;
  ka.out0 = out_c0;
  ka.out1 = out_c1;
  ka.out_stride = poly_words;
  ka.accumulate = accumulate;
  ProfScope prof(ctx, "keyswitch");
  return shg::launch_keyswitch(ka, gd, v & 3, ctx->d_primes, lp, count, num_limbs, log_n, s);
}

// Digit planes: the inverse transform writes the gadget digits as 16-bit
// planes read straight from the components (FP64 key switch, bases up to
// 2^16, NTT-form source).
static bool keyswitch_planes_ok(int v, int log_n, const shg::GadgetDesc& gd) {
  return log_n >= 4 && log_n <= shg::kMaxLogM && (v & 3) == SHG_NTT_F64 && gd.small_digits &&
         gd.max_log_b <= 16;
}

// Shared body of the two key-switch entry points: for each component index
// in `srcs`, bring it to coefficient form (as digit planes or an exported
// copy) and run the fused key-switch kernel (first call initialises from
// c0/c1, later ones accumulate).  The caller holds ctx->mu: the work buffers
// are the context's.
static int keyswitch_locked(shg_context* ctx, const uint64_t* comps, int comps_per_ct,
                            const int* src_comp, const int* src_is_ntt, const int* key_base,
                            int num_src, bool init_c1, const uint64_t* key_as,
                            const uint64_t* key_bs, const int* log_bs, uint64_t* out_c0,
                            uint64_t* out_c1, int count, int num_limbs, const int* primes,
                            const LimbPrimes& lp, int log_n, const shg::GadgetDesc& gd,
                            const uint64_t* f64_as = nullptr, const uint64_t* f64_bs = nullptr) {
  const size_t n = (size_t)1 << log_n;
  const size_t poly_words = (size_t)num_limbs * n;
  int rc = SHG_OK;
  uint64_t* src = nullptr;  // coefficient-form copy of a source component (scratch)
  hipStream_t s = ctx->stream;
  const int v = resolve_variant(ctx, SHG_NTT_AUTO, primes, num_limbs, log_n,
                                (long long)count * num_limbs);
  hipError_t e = hipSuccess;
  const bool fused = log_n <= shg::kMaxLogM && log_n >= 4;
  const bool planes_ok = keyswitch_planes_ok(v, log_n, gd);
  uint16_t* planes = nullptr;
  if (planes_ok) {
    rc = ensure_buf(ctx, ctx->ks_planes, (size_t)count * gd.k * n * 2);
    if (rc) return rc;
    planes = (uint16_t*)ctx->ks_planes.p;
  }
  for (int i = 0; i < num_src && e == hipSuccess; ++i) {
    const bool use_planes = planes_ok && src_is_ntt[i];
    if (!use_planes) {
      rc = ensure_scratch(ctx, (size_t)count * poly_words * 8);
      if (rc) return rc;
      src = (uint64_t*)ctx->scratch;
    }
    if (fused) {
      e = keyswitch_fused_one(ctx, v, use_planes ? planes : nullptr, src, comps, comps_per_ct,
                              src_comp[i], src_is_ntt[i] != 0, key_base[i], init_c1, i > 0,
                              key_as, key_bs, out_c0, out_c1, count, num_limbs, lp, log_n, gd,
                              f64_as, f64_bs);
      continue;
    }
    // The unfused path decomposes Montgomery representations of the
    // coefficient-form source.
    e = hipMemcpy2DAsync(src, poly_words * 8, comps + (size_t)src_comp[i] * poly_words,
                         (size_t)comps_per_ct * poly_words * 8, poly_words * 8, count,
                         hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && src_is_ntt[i]) {
      ProfScope prof(ctx, "keyswitch_intt");
      e = shg::launch_ntt(true, v, src, ctx->d_primes, lp, count * num_limbs, num_limbs, log_n,
                          shg::kMaxLogM, s, 1);
    }
    if (e != hipSuccess) break;
    {
      // Unfused form for ring degrees outside the fused kernel's range:
      // materialise the digits, transform them, and multiply-accumulate.
      const size_t need = (size_t)count * poly_words * 8 * (1 + (size_t)gd.k);
      if (need > ctx->scratch_size) {
        // Grow while keeping the coefficient-form source.
        if (capturing(ctx)) return grow_refused(need);
        void* bigger = nullptr;
        if (hipMalloc(&bigger, need) != hipSuccess)
          return fail(SHG_E_NO_MEMORY, "scratch allocation of %zu bytes failed", need);
        hipError_t ge =
            hipMemcpyAsync(bigger, src, (size_t)count * poly_words * 8, hipMemcpyDeviceToDevice, s);
        if (ge != hipSuccess) {
          (void)hipFree(bigger);  // already failing; report the copy
          return hip_fail(ge, "scratch growth");
        }
        rc = release_work_buffer(ctx, ctx->scratch);
        if (rc) {
          (void)hipFree(bigger);  // already failing; report the release
          return rc;
        }
        ctx->scratch = bigger;
        ctx->scratch_size = need;
        src = (uint64_t*)bigger;
      }
      uint64_t* digits = src + (size_t)count * poly_words;
      e = shg::launch_gadget_decompose(src, digits, gd, ctx->d_primes, lp, count, num_limbs,
                                       log_n, s);
      if (e == hipSuccess)
        e = shg::launch_ntt(false, v, digits, ctx->d_primes, lp, count * gd.k * num_limbs,
                            num_limbs, log_n, shg::kMaxLogM, s);
      if (e == hipSuccess && i == 0) {
        e = hipMemcpy2DAsync(out_c0, poly_words * 8, comps, (size_t)comps_per_ct * poly_words * 8,
                             poly_words * 8, count, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) {
          if (init_c1)
            e = hipMemcpy2DAsync(out_c1, poly_words * 8, comps + poly_words,
                                 (size_t)comps_per_ct * poly_words * 8, poly_words * 8, count,
                                 hipMemcpyDeviceToDevice, s);
          else
            e = shg::launch_zero(out_c1, (size_t)count * poly_words, s);
        }
      }
      for (int j = 0; j < gd.k && e == hipSuccess; ++j) {
        const uint64_t* dj = digits + (size_t)j * poly_words;
        e = shg::launch_pointwise(SHG_OP_FMA, out_c0, dj, (size_t)gd.k * poly_words,
                                  key_bs + (size_t)(key_base[i] + j) * poly_words, 0,
                                  ctx->d_primes, lp, count, num_limbs, log_n, s);
        if (e == hipSuccess)
          e = shg::launch_pointwise(SHG_OP_FMA, out_c1, dj, (size_t)gd.k * poly_words,
                                    key_as + (size_t)(key_base[i] + j) * poly_words, 0,
                                    ctx->d_primes, lp, count, num_limbs, log_n, s);
      }
    }
  }
  (void)log_bs;
  return e == hipSuccess ? SHG_OK : hip_fail(e, "key switch launch");
}

static int keyswitch_common(shg_context* ctx, const uint64_t* comps, int comps_per_ct,
                            const int* src_comp, const int* src_is_ntt, const int* key_base,
                            int num_src, bool init_c1, const uint64_t* key_as,
                            const uint64_t* key_bs, const int* log_bs, uint64_t* out_c0,
                            uint64_t* out_c1, int count, int num_limbs, const int* primes,
                            const LimbPrimes& lp, int log_n, const shg::GadgetDesc& gd,
                            const uint64_t* f64_as = nullptr, const uint64_t* f64_bs = nullptr) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  return keyswitch_locked(ctx, comps, comps_per_ct, src_comp, src_is_ntt, key_base, num_src,
                          init_c1, key_as, key_bs, log_bs, out_c0, out_c1, count, num_limbs, primes,
                          lp, log_n, gd, f64_as, f64_bs);
}

}  // extern "C"

// Argument checks of RnsRelinKey::ApplyToRlweCiphertext
// (rns/rns_relinearization_key.cc:184-232) shared by both relinearization
// entry points; degree 1 is a plain copy (nothing to switch).  *done is set
// when the call is complete (an error, count 0, or the copy).
static int relin_prepare(shg_context* ctx, const uint64_t* comps, int degree, int ntt_mask,
                         const uint64_t* key_as, const uint64_t* key_bs, int num_key_polys,
                         const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1, int count,
                         int num_limbs, const int* primes, LimbPrimes* lp, int* log_n,
                         shg::GadgetDesc* gd, bool* done) {
  *done = true;
  int rc = check_primes(ctx, primes, num_limbs, log_n, lp);
  if (rc) return rc;
  rc = build_gadget(ctx, primes, num_limbs, host_log_bs, gd);
  if (rc) return rc;
  if (degree < 1 || degree > 16)
    return fail(SHG_E_INVALID_ARGUMENT, "ciphertext degree must be in [1, 16].");
  if (!(ntt_mask & 1)) return fail(SHG_E_INVALID_ARGUMENT, "RNS polynomial `this` must be in NTT form.");
  if (!(ntt_mask & 2)) return fail(SHG_E_INVALID_ARGUMENT, "RNS polynomial `this` must be in NTT form.");
  const int needed = degree >= 2 ? (degree - 2) * degree + gd->k : 0;
  if (num_key_polys < needed)
    return fail(SHG_E_INVALID_ARGUMENT,
                "`ciphertext` degree is larger than degree of this relinearization key.");
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned16(comps) || !aligned16(out_c0) || !aligned16(out_c1) ||
      (degree >= 2 && (!aligned16(key_as) || !aligned16(key_bs))))
    return fail(SHG_E_INVALID_ARGUMENT, "data must be non-null 16-byte aligned pointers.");
  const size_t poly_words = (size_t)num_limbs << *log_n;
  if (degree == 1) {
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemcpy2DAsync(out_c0, poly_words * 8, comps, 2 * poly_words * 8,
                                    poly_words * 8, count, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(out_c1, poly_words * 8, comps + poly_words, 2 * poly_words * 8,
                           poly_words * 8, count, hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? SHG_OK : hip_fail(e, "copy");
  }
  *done = false;
  return SHG_OK;
}

extern "C" {

static int relin_apply_impl(shg_context* ctx, const uint64_t* comps, int degree, int ntt_mask,
                            const uint64_t* key_as, const uint64_t* key_bs, int num_key_polys,
                            const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1,
                            int count, int num_limbs, const int* primes, const uint64_t* f64_as,
                            const uint64_t* f64_bs);

int shg_relin_apply(shg_context* ctx, const uint64_t* comps, int degree, int ntt_mask,
                    const uint64_t* key_as, const uint64_t* key_bs, int num_key_polys,
                    const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1, int count,
                    int num_limbs, const int* primes) {
  return relin_apply_impl(ctx, comps, degree, ntt_mask, key_as, key_bs, num_key_polys,
                          host_log_bs, out_c0, out_c1, count, num_limbs, primes, nullptr, nullptr);
}

static int relin_apply_impl(shg_context* ctx, const uint64_t* comps, int degree, int ntt_mask,
                            const uint64_t* key_as, const uint64_t* key_bs, int num_key_polys,
                            const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1,
                            int count, int num_limbs, const int* primes, const uint64_t* f64_as,
                            const uint64_t* f64_bs) {
  LimbPrimes lp{};
  int log_n = 0;
  shg::GadgetDesc gd{};
  bool done = true;
  int rc = relin_prepare(ctx, comps, degree, ntt_mask, key_as, key_bs, num_key_polys, host_log_bs,
                         out_c0, out_c1, count, num_limbs, primes, &lp, &log_n, &gd, &done);
  if (rc || done) return rc;
  int src_comp[16], src_ntt[16], base[16];
  for (int i = 2; i <= degree; ++i) {
    src_comp[i - 2] = i;
    src_ntt[i - 2] = (ntt_mask >> i) & 1;
    base[i - 2] = (i - 2) * degree;
  }
  return keyswitch_common(ctx, comps, degree + 1, src_comp, src_ntt, base, degree - 1, true,
                          key_as, key_bs, host_log_bs, out_c0, out_c1, count, num_limbs, primes,
                          lp, log_n, gd, f64_as, f64_bs);
}

int shg_relin_apply_noalloc(shg_context* ctx, const uint64_t* comps, int degree, int ntt_mask,
                            const uint64_t* key_as, const uint64_t* key_bs, int num_key_polys,
                            const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1,
                            int count, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  shg::GadgetDesc gd{};
  bool done = true;
  int rc = relin_prepare(ctx, comps, degree, ntt_mask, key_as, key_bs, num_key_polys, host_log_bs,
                         out_c0, out_c1, count, num_limbs, primes, &lp, &log_n, &gd, &done);
  if (rc || done) return rc;
  if (log_n < 4 || log_n > shg::kMaxLogM)
    return fail(SHG_E_UNSUPPORTED,
                "shg_relin_apply_noalloc supports ring degrees 2^4 to 2^%d; use shg_relin_apply.",
                shg::kMaxLogM);
  if (!ctx->noalloc_work)
    return fail(SHG_E_DEVICE, "the no-allocation work area of device %d is unavailable: %s",
                ctx->device, ctx->noalloc_error.c_str());
  std::lock_guard<std::mutex> lock(ctx->mu);  // the autotuner's table
  const int v = resolve_variant(ctx, SHG_NTT_AUTO, primes, num_limbs, log_n,
                                (long long)count * num_limbs);
  const bool planes_ok = keyswitch_planes_ok(v, log_n, gd);
  const size_t n = (size_t)1 << log_n;
  const size_t poly_words = (size_t)num_limbs * n;
  // Work space per ciphertext: the digit planes of an NTT-form source, or
  // the exported copy of a source that takes the other path.  Sources run
  // one after the other on the stream and share the area.
  size_t per_ct = 0;
  for (int i = 2; i <= degree; ++i) {
    const bool use_planes = planes_ok && ((ntt_mask >> i) & 1);
    const size_t need = use_planes ? (size_t)gd.k * n * 2 : poly_words * 8;
    per_ct = need > per_ct ? need : per_ct;
  }
  if (per_ct > kNoAllocWorkBytes)
    return fail(SHG_E_INVALID_ARGUMENT,
                "one ciphertext needs %zu bytes of work space, above the %zu bytes the "
                "no-allocation key switch holds per device.", per_ct, kNoAllocWorkBytes);
  // Ciphertexts are switched in chunks that fit the work area.
  const int chunk = (int)std::min<size_t>((size_t)count, kNoAllocWorkBytes / per_ct);
  const size_t ct_words = (size_t)(degree + 1) * poly_words;
  hipError_t e = hipSuccess;
  for (int c0 = 0; c0 < count && e == hipSuccess; c0 += chunk) {
    const int cnt = std::min(chunk, count - c0);
    for (int i = 2; i <= degree && e == hipSuccess; ++i) {
      const bool is_ntt = (ntt_mask >> i) & 1;
      const bool use_planes = planes_ok && is_ntt;
      e = keyswitch_fused_one(ctx, v, use_planes ? (uint16_t*)ctx->noalloc_work : nullptr,
                              use_planes ? nullptr : (uint64_t*)ctx->noalloc_work,
                              comps + (size_t)c0 * ct_words, degree + 1, i, is_ntt,
                              (i - 2) * degree, true, i > 2, key_as, key_bs,
                              out_c0 + (size_t)c0 * poly_words, out_c1 + (size_t)c0 * poly_words,
                              cnt, num_limbs, lp, log_n, gd);
    }
  }
  return e == hipSuccess ? SHG_OK : hip_fail(e, "key switch launch");
}

static int galois_apply_impl(shg_context* ctx, const uint64_t* comps, int c1_is_ntt,
                             const uint64_t* key_as, const uint64_t* key_bs,
                             const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1,
                             int count, int num_limbs, const int* primes,
                             const uint64_t* f64_as, const uint64_t* f64_bs);

int shg_galois_apply(shg_context* ctx, const uint64_t* comps, int c1_is_ntt, const uint64_t* key_as,
                     const uint64_t* key_bs, const int* host_log_bs, uint64_t* out_c0,
                     uint64_t* out_c1, int count, int num_limbs, const int* primes) {
  return galois_apply_impl(ctx, comps, c1_is_ntt, key_as, key_bs, host_log_bs, out_c0, out_c1,
                           count, num_limbs, primes, nullptr, nullptr);
}

static int galois_apply_impl(shg_context* ctx, const uint64_t* comps, int c1_is_ntt,
                             const uint64_t* key_as, const uint64_t* key_bs,
                             const int* host_log_bs, uint64_t* out_c0, uint64_t* out_c1,
                             int count, int num_limbs, const int* primes,
                             const uint64_t* f64_as, const uint64_t* f64_bs) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  shg::GadgetDesc gd{};
  rc = build_gadget(ctx, primes, num_limbs, host_log_bs, &gd);
  if (rc) return rc;
  if (count <= 0) return count == 0 ? SHG_OK : fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (!aligned16(comps) || !aligned16(out_c0) || !aligned16(out_c1) || !aligned16(key_as) ||
      !aligned16(key_bs))
    return fail(SHG_E_INVALID_ARGUMENT, "data must be non-null 16-byte aligned pointers.");
  const int src_comp = 1, src_ntt = c1_is_ntt ? 1 : 0, base = 0;
  return keyswitch_common(ctx, comps, 2, &src_comp, &src_ntt, &base, 1, false, key_as, key_bs,
                          host_log_bs, out_c0, out_c1, count, num_limbs, primes, lp, log_n, gd,
                          f64_as, f64_bs);
}

// A key-switching key with FP64-form copies of its residues for the
// digit-plane kernel; the Montgomery words stay the caller's.
struct shg_ks_key {
  const uint64_t* key_as = nullptr;
  const uint64_t* key_bs = nullptr;
  uint64_t* f64_as = nullptr;  // owned; null when a modulus is 2^50 or more
  uint64_t* f64_bs = nullptr;
  int num_key_polys = 0;
  int num_limbs = 0;
  int primes[SHG_MAX_LIMBS] = {};
};

int shg_ks_key_create(shg_context* ctx, const uint64_t* key_as, const uint64_t* key_bs,
                      int num_key_polys, int num_limbs, const int* primes, shg_ks_key** out) {
  if (!ctx || !out) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (num_key_polys <= 0) return fail(SHG_E_INVALID_ARGUMENT, "num_key_polys must be positive.");
  if (!aligned16(key_as) || !aligned16(key_bs))
    return fail(SHG_E_INVALID_ARGUMENT, "keys must be non-null 16-byte aligned pointers.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_ks_key_create");
  auto key = std::make_unique<shg_ks_key>();
  key->key_as = key_as;
  key->key_bs = key_bs;
  key->num_key_polys = num_key_polys;
  key->num_limbs = num_limbs;
  for (int i = 0; i < num_limbs; ++i) key->primes[i] = primes[i];
  if (!width_class(ctx, primes, num_limbs)) {
    const size_t words = (size_t)num_key_polys * num_limbs << log_n;
    if (hipMalloc((void**)&key->f64_as, words * 8) != hipSuccess)
      return fail(SHG_E_NO_MEMORY, "key copy of %zu bytes failed.", words * 8);
    if (hipMalloc((void**)&key->f64_bs, words * 8) != hipSuccess) {
      (void)hipFree(key->f64_as);  // already failing; report the allocation
      return fail(SHG_E_NO_MEMORY, "key copy of %zu bytes failed.", words * 8);
    }
    hipError_t e = shg::launch_key_to_f64(key_as, key->f64_as, ctx->d_primes, lp,
                                          num_key_polys, num_limbs, log_n, ctx->stream);
    if (e == hipSuccess)
      e = shg::launch_key_to_f64(key_bs, key->f64_bs, ctx->d_primes, lp, num_key_polys,
                                 num_limbs, log_n, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      (void)hipFree(key->f64_as);  // already failing; report the launch
      (void)hipFree(key->f64_bs);
      return hip_fail(e, "key conversion");
    }
  }
  *out = key.release();
  return SHG_OK;
}

int shg_ks_key_destroy(shg_context* ctx, shg_ks_key* key) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (!key) return SHG_OK;
  if (capturing(ctx)) return refuse_capture(ctx, "shg_ks_key_destroy");
  (void)hipStreamSynchronize(ctx->stream);  // launches that read the copies have finished
  if (key->f64_as) (void)hipFree(key->f64_as);
  if (key->f64_bs) (void)hipFree(key->f64_bs);
  delete key;
  return SHG_OK;
}

static int check_key(const shg_ks_key* key, int num_limbs, const int* primes) {
  if (!key) return fail(SHG_E_INVALID_ARGUMENT, "`key` must not be null.");
  if (!primes || key->num_limbs != num_limbs)
    return fail(SHG_E_INVALID_ARGUMENT, "`key` was made for %d RNS moduli.", key->num_limbs);
  for (int i = 0; i < num_limbs; ++i)
    if (key->primes[i] != primes[i])
      return fail(SHG_E_INVALID_ARGUMENT, "`key` was made for other RNS moduli.");
  return SHG_OK;
}

int shg_relin_apply_key(shg_context* ctx, const uint64_t* comps, int degree, int ntt_mask,
                        const shg_ks_key* key, const int* host_log_bs, uint64_t* out_c0,
                        uint64_t* out_c1, int count, int num_limbs, const int* primes) {
  int rc = check_key(key, num_limbs, primes);
  if (rc) return rc;
  return relin_apply_impl(ctx, comps, degree, ntt_mask, key->key_as, key->key_bs,
                          key->num_key_polys, host_log_bs, out_c0, out_c1, count, num_limbs,
                          primes, key->f64_as, key->f64_bs);
}

int shg_galois_apply_key(shg_context* ctx, const uint64_t* comps, int c1_is_ntt,
                         const shg_ks_key* key, const int* host_log_bs, uint64_t* out_c0,
                         uint64_t* out_c1, int count, int num_limbs, const int* primes) {
  int rc = check_key(key, num_limbs, primes);
  if (rc) return rc;
  return galois_apply_impl(ctx, comps, c1_is_ntt, key->key_as, key->key_bs, host_log_bs, out_c0,
                           out_c1, count, num_limbs, primes, key->f64_as, key->f64_bs);
}

// The NTT-form monomial -X^(n - 2^level) = X^(-2^level) mod each prime
// (the to_absorb of oblivious_expand.h:195-200), built once per prime set.
static int expand_monomial(shg_context* ctx, const int* primes, int num_limbs, int log_n, int level,
                           const LimbPrimes& lp, const uint64_t** out) {
  std::vector<int> key(primes, primes + num_limbs);
  auto it = ctx->monomials.find({key, level});
  if (it != ctx->monomials.end()) {
    *out = (const uint64_t*)it->second;
    return SHG_OK;
  }
  const size_t n = (size_t)1 << log_n;
  std::vector<uint64_t> host((size_t)num_limbs * n, 0);
  for (int i = 0; i < num_limbs; ++i) {
    const Params* P = ctx->primes[primes[i]]->mp.get();
    host[(size_t)i * n + (n - ((size_t)1 << level))] =
        ModInt::ImportInt(P->modulus - 1, P).value().GetMontgomeryRepresentation();
  }
  void* d = nullptr;
  if (hipMalloc(&d, host.size() * 8) != hipSuccess)
    return fail(SHG_E_NO_MEMORY, "monomial allocation failed.");
  hipError_t e = hipMemcpy(d, host.data(), host.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = shg::launch_ntt(false, resolve_variant(ctx, SHG_NTT_AUTO, primes, num_limbs, log_n, num_limbs),
                        (uint64_t*)d, ctx->d_primes, lp, num_limbs, num_limbs, log_n, shg::kMaxLogM,
                        ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(d);  // already failing; report the first error
    return hip_fail(e, "monomial");
  }
  ctx->monomials[{key, level}] = d;
  *out = (const uint64_t*)d;
  return SHG_OK;
}

static int expand_level_impl(shg_context* ctx, uint64_t* const* host_cts, int count, int level,
                             const uint64_t* key_as, const uint64_t* key_bs,
                             const int* host_log_bs, int num_limbs, const int* primes,
                             const uint64_t* f64_as, const uint64_t* f64_bs);

int shg_oblivious_expand_level(shg_context* ctx, uint64_t* const* host_cts, int count, int level,
                               const uint64_t* key_as, const uint64_t* key_bs,
                               const int* host_log_bs, int num_limbs, const int* primes) {
  return expand_level_impl(ctx, host_cts, count, level, key_as, key_bs, host_log_bs, num_limbs,
                           primes, nullptr, nullptr);
}

int shg_oblivious_expand_level_key(shg_context* ctx, uint64_t* const* host_cts, int count,
                                   int level, const shg_ks_key* key, const int* host_log_bs,
                                   int num_limbs, const int* primes) {
  int rc = check_key(key, num_limbs, primes);
  if (rc) return rc;
  return expand_level_impl(ctx, host_cts, count, level, key->key_as, key->key_bs, host_log_bs,
                           num_limbs, primes, key->f64_as, key->f64_bs);
}

static int expand_level_impl(shg_context* ctx, uint64_t* const* host_cts, int count, int level,
                             const uint64_t* key_as, const uint64_t* key_bs,
                             const int* host_log_bs, int num_limbs, const int* primes,
                             const uint64_t* f64_as, const uint64_t* f64_bs) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (count < 0 || (count > 0 && !host_cts))
    return fail(SHG_E_INVALID_ARGUMENT, "queries must be a non-negative count with pointers.");
  if (level < 0 || level >= log_n)
    return fail(SHG_E_INVALID_ARGUMENT, "level must be in [0, log_n).");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_oblivious_expand_level");
  if (log_n < 4 || log_n > shg::kMaxLogM)
    return fail(SHG_E_INVALID_ARGUMENT, "expansion needs 2^4 <= n <= 2^14.");
  shg::GadgetDesc gd{};
  rc = build_gadget(ctx, primes, num_limbs, host_log_bs, &gd);
  if (rc) return rc;
  if (count == 0) return SHG_OK;
  if (!aligned16(key_as) || !aligned16(key_bs))
    return fail(SHG_E_INVALID_ARGUMENT, "keys must be non-null 16-byte aligned pointers.");
  for (int g = 0; g < count; ++g)
    if (!aligned16(host_cts[g]))
      return fail(SHG_E_INVALID_ARGUMENT, "query buffers must be 16-byte aligned device memory.");
  const int m = 1 << level;
  const int power = (1 << (log_n - level)) + 1;  // n / 2^j + 1
  const size_t poly = (size_t)num_limbs << log_n;
  const size_t cts = (size_t)count * m;
  const uint64_t* mono = nullptr;
  uint64_t *S = nullptr, *sub0 = nullptr, *sub1 = nullptr;
  uint64_t** d_cts = nullptr;
  // The whole level holds the context's lock: its three launches share the
  // context's work buffers (S, sub0/sub1 and the pointer table).
  std::lock_guard<std::mutex> lock(ctx->mu);
  {
    rc = expand_monomial(ctx, primes, num_limbs, log_n, level, lp, &mono);
    if (!rc) rc = ensure_buf(ctx, ctx->exp_s, cts * 2 * poly * 8);
    if (!rc) rc = ensure_buf(ctx, ctx->exp_sub, cts * 2 * poly * 8);
    if (!rc) rc = ensure_buf(ctx, ctx->exp_ptrs, (size_t)count * sizeof(uint64_t*));
    if (rc) return rc;
    S = (uint64_t*)ctx->exp_s.p;
    sub0 = (uint64_t*)ctx->exp_sub.p;
    sub1 = sub0 + cts * poly;
    d_cts = (uint64_t**)ctx->exp_ptrs.p;
    hipError_t e = hipMemcpyAsync(d_cts, host_cts, (size_t)count * sizeof(uint64_t*),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = shg::launch_expand_substitute(d_cts, count, m, S, power, log_n, num_limbs, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "expansion substitute");
  }
  const int src_comp = 1, src_ntt = 1, base = 0;
  rc = keyswitch_locked(ctx, S, 2, &src_comp, &src_ntt, &base, 1, false, key_as, key_bs,
                        host_log_bs, sub0, sub1, (int)cts, num_limbs, primes, lp, log_n, gd,
                        f64_as, f64_bs);
  if (rc) return rc;
  hipError_t e = shg::launch_expand_update(d_cts, count, m, sub0, sub1, mono, ctx->d_primes, lp,
                                           log_n, num_limbs, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "expansion update");
}

int shg_oblivious_expand(shg_context* ctx, uint64_t* const* host_cts, int count, int levels,
                         const uint64_t* const* key_as, const uint64_t* const* key_bs,
                         const int* host_log_bs, int num_limbs, const int* primes) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (levels < 0 || (levels > 0 && (!key_as || !key_bs)))
    return fail(SHG_E_INVALID_ARGUMENT, "one Galois key per level is required.");
  for (int j = 0; j < levels; ++j) {
    int rc = shg_oblivious_expand_level(ctx, host_cts, count, j, key_as[j], key_bs[j], host_log_bs,
                                        num_limbs, primes);
    if (rc) return rc;
  }
  return SHG_OK;
}

int shg_pir_db_export(shg_context* ctx, const uint64_t* db, long long num_records, uint64_t* out,
                      int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (num_records < 0) return fail(SHG_E_INVALID_ARGUMENT, "num_records must be non-negative.");
  if (num_records == 0) return SHG_OK;
  if (!aligned_ptr(db) || !aligned_ptr(out)) return fail(SHG_E_INVALID_ARGUMENT, "data must not be null.");
  hipError_t e = shg::launch_export_words(db, out, num_records, ctx->d_primes, lp, log_n, num_limbs,
                                          ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "export");
}

int shg_pir_db_synthesize(shg_context* ctx, uint64_t seed, uint64_t t, long long num_records,
                          uint64_t* out, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (t == 0) return fail(SHG_E_INVALID_ARGUMENT, "plaintext modulus must be positive.");
  if (num_records < 0 || num_records * num_limbs > 0x7fffffffLL)
    return fail(SHG_E_INVALID_ARGUMENT, "num_records out of range.");
  if (num_records == 0) return SHG_OK;
  if (!aligned16(out)) return fail(SHG_E_INVALID_ARGUMENT, "out must be 16-byte aligned.");
  hipError_t e = shg::launch_db_synth(out, num_records, seed, t, ctx->d_primes, lp, log_n,
                                      num_limbs, ctx->stream);
  if (e == hipSuccess)
    e = shg::launch_ntt(false, resolve_variant(ctx, SHG_NTT_AUTO, primes, num_limbs, log_n,
                                               num_records * num_limbs),
                        out, ctx->d_primes, lp, (int)(num_records * num_limbs), num_limbs, log_n,
                        shg::kMaxLogM, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "database synthesis");
}

int shg_pir_inner_product(shg_context* ctx, const uint64_t* const* host_expanded, int count,
                          int levels, const uint64_t* db_exported, int num_columns,
                          uint64_t* const* host_out, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (count < 0 || (count > 0 && (!host_expanded || !host_out)))
    return fail(SHG_E_INVALID_ARGUMENT, "queries must be a non-negative count with pointers.");
  if (levels < 0 || levels > 20 || num_columns <= 0)
    return fail(SHG_E_INVALID_ARGUMENT, "levels must be in [0, 20] and num_columns positive.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_pir_inner_product");
  if (count == 0) return SHG_OK;
  if (!aligned_ptr(db_exported)) return fail(SHG_E_INVALID_ARGUMENT, "database must not be null.");
  for (int g = 0; g < count; ++g)
    if (!aligned_ptr(host_expanded[g]) || !aligned_ptr(host_out[g]))
      return fail(SHG_E_INVALID_ARGUMENT, "query buffers must not be null.");
  for (int i = 0; i < num_limbs; ++i)
    if (ctx->primes[primes[i]]->q >= (1ull << 50))
      return fail(SHG_E_INVALID_ARGUMENT, "the inner product requires moduli below 2^50.");
  const int K = 1 << levels;
  const size_t poly = (size_t)num_limbs << log_n;
  const int ks = shg::pir_k_split(count, num_columns, K, log_n, num_limbs);
  std::lock_guard<std::mutex> lock(ctx->mu);
  rc = ensure_buf(ctx, ctx->pir_part, (size_t)count * num_columns * ks * 2 * poly * 8);
  if (!rc) rc = ensure_buf(ctx, ctx->exp_ptrs2, (size_t)count * 2 * sizeof(uint64_t*));
  if (rc) return rc;
  std::vector<const uint64_t*> ptrs(2 * (size_t)count);
  for (int g = 0; g < count; ++g) {
    ptrs[g] = host_expanded[g];
    ptrs[count + g] = host_out[g];
  }
  auto** d_ptrs = (const uint64_t**)ctx->exp_ptrs2.p;
  hipError_t e = hipMemcpyAsync(d_ptrs, ptrs.data(), ptrs.size() * sizeof(uint64_t*),
                                hipMemcpyHostToDevice, ctx->stream);
  ProfScope prof(ctx, "pir_inner_product");
  if (e == hipSuccess)
    e = shg::launch_pir_inner_product(d_ptrs, host_expanded[0], count, K, db_exported, num_columns,
                                      (uint64_t*)ctx->pir_part.p, ks, (uint64_t* const*)(d_ptrs + count),
                                      ctx->d_primes, lp, log_n, num_limbs, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "inner product");
}

int shg_lazy_fma(shg_context* ctx, const uint64_t* as, size_t as_stride, const uint64_t* bs,
                 size_t bs_stride, int num_terms, uint64_t* out, int num_limbs, const int* primes) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (num_terms < 0) return fail(SHG_E_INVALID_ARGUMENT, "num_terms must be non-negative.");
  if (!aligned_ptr(out) || (num_terms > 0 && (!aligned_ptr(as) || !aligned_ptr(bs))))
    return fail(SHG_E_INVALID_ARGUMENT, "data must not be null.");
  for (int i = 0; i < num_limbs; ++i)
    if (127 - 2 * (int)ctx->primes[primes[i]]->mp->log_modulus < 1)
      return fail(SHG_E_INVALID_ARGUMENT,
                  "The RNS moduli do not allow for lazy polynomials, as the logarithm of the "
                  "maximum level is stricly smaller than 1.");
  hipError_t e = shg::launch_lazy_fma(as, as_stride, bs, bs_stride, num_terms, out, ctx->d_primes,
                                      lp, num_limbs, log_n, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "lazy fma launch");
}

// The serialized layout of one polynomial over `primes`
// (MontgomeryInt::SerializeVector per limb, montgomery.cc:143-167).
static int wire_desc(shg_context* ctx, const int* primes, int num_limbs, shg::WireDesc* wd) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  const long long n = 1LL << log_n;
  if (n > (1LL << 15))
    return fail(SHG_E_INVALID_ARGUMENT,
                "Number of coefficients, %lld, cannot be larger than 32768.", n);
  wd->num_limbs = num_limbs;
  wd->log_n = log_n;
  wd->aligned = 1;
  unsigned long long off = 0;
  for (int i = 0; i < num_limbs; ++i) {
    const int b = (int)ctx->primes[primes[i]]->mp->log_modulus;
    wd->bits[i] = (unsigned char)b;
    wd->limb_off[i] = off;
    const unsigned long long bytes = ((unsigned long long)n * b + 7) / 8;
    if (bytes % 8) wd->aligned = 0;
    off += bytes;
  }
  wd->limb_off[num_limbs] = off;
  return SHG_OK;
}

int shg_serialized_bytes(shg_context* ctx, const int* primes, int num_limbs, size_t* limb_bytes,
                         size_t* poly_bytes) {
  shg::WireDesc wd{};
  int rc = wire_desc(ctx, primes, num_limbs, &wd);
  if (rc) return rc;
  if (limb_bytes)
    for (int i = 0; i < num_limbs; ++i) limb_bytes[i] = wd.limb_off[i + 1] - wd.limb_off[i];
  if (poly_bytes) *poly_bytes = wd.limb_off[num_limbs];
  return SHG_OK;
}

int shg_serialize_polys(shg_context* ctx, const uint64_t* polys, int count, int num_limbs,
                        const int* primes, uint8_t* out) {
  shg::WireDesc wd{};
  int rc = wire_desc(ctx, primes, num_limbs, &wd);
  if (rc) return rc;
  if (count < 0) return fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (count == 0) return SHG_OK;
  if (!aligned_ptr(polys) || !out) return fail(SHG_E_INVALID_ARGUMENT, "data must not be null.");
  if ((uintptr_t)out % 8) wd.aligned = 0;
  hipError_t e = shg::launch_pack(polys, out, wd, count, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "serialize launch");
}

int shg_deserialize_polys(shg_context* ctx, const uint8_t* in, int count, int num_limbs,
                          const int* primes, uint64_t* polys) {
  shg::WireDesc wd{};
  int rc = wire_desc(ctx, primes, num_limbs, &wd);
  if (rc) return rc;
  if (count < 0) return fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (count == 0) return SHG_OK;
  if (!aligned_ptr(polys) || !in) return fail(SHG_E_INVALID_ARGUMENT, "data must not be null.");
  if ((uintptr_t)in % 8) wd.aligned = 0;
  hipError_t e = shg::launch_unpack(in, polys, wd, count, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "deserialize launch");
}

int shg_sample_uniform_chacha(shg_context* ctx, const uint8_t* host_seed, size_t seed_len,
                              int count, int num_limbs, const int* primes, uint64_t* out) {
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (!host_seed || seed_len != 32)
    return fail(SHG_E_INVALID_ARGUMENT,
                "Cannot create Prng with key of the wrong size. Real key length of %zu instead "
                "of expected key length of 32.", host_seed ? seed_len : (size_t)0);
  if (count < 0) return fail(SHG_E_INVALID_ARGUMENT, "count must be non-negative.");
  if (count == 0) return SHG_OK;
  if (!aligned_ptr(out)) return fail(SHG_E_INVALID_ARGUMENT, "out must not be null.");
  shg::PadDesc pd{};
  for (int i = 0; i < 8; ++i)
    pd.key[i] = (uint32_t)host_seed[4 * i] | (uint32_t)host_seed[4 * i + 1] << 8 |
                (uint32_t)host_seed[4 * i + 2] << 16 | (uint32_t)host_seed[4 * i + 3] << 24;
  pd.count = count;
  pd.num_limbs = num_limbs;
  pd.log_n = log_n;
  for (int i = 0; i < num_limbs; ++i) {
    const Params* P = ctx->primes[primes[i]]->mp.get();
    if (P->log_modulus <= 8)
      return fail(SHG_E_UNSUPPORTED, "moduli of 8 bits or fewer are not supported here.");
    pd.bits[i] = (unsigned char)P->log_modulus;
    pd.q[i] = P->modulus;
  }
  hipError_t e = shg::launch_chacha_pad(pd, out, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "sample launch");
}

int shg_profile_enable(shg_context* ctx, int enable) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_profile_enable");
  std::lock_guard<std::mutex> lock(ctx->mu);
  prof_collect(ctx);
  ctx->profiling = enable != 0;
  return SHG_OK;
}

int shg_profile_reset(shg_context* ctx) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_profile_reset");
  std::lock_guard<std::mutex> lock(ctx->mu);
  prof_collect(ctx);
  ctx->prof_totals.clear();
  return SHG_OK;
}

int shg_profile_query(shg_context* ctx, const char* kernel, double* total_ms, long long* launches) {
  if (!ctx || !kernel || !total_ms || !launches)
    return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_profile_query");
  std::lock_guard<std::mutex> lock(ctx->mu);
  prof_collect(ctx);
  auto it = ctx->prof_totals.find(kernel);
  *total_ms = it == ctx->prof_totals.end() ? 0.0 : it->second.first;
  *launches = it == ctx->prof_totals.end() ? 0 : it->second.second;
  return SHG_OK;
}

}  // extern "C"

struct shg_batch_plan {
  void* d_jobs = nullptr;  // LimbJobDev[] or PolyJobDev[]
  bool modreduce = false;
  int num_jobs = 0;
  int variant = SHG_NTT_F64;
  int grid = 0;
  size_t shmem = 0;
  bool has_other = true;     // every job has a second operand (multiply allowed)
  uint64_t* scratch = nullptr;  // modulus switch: 2^16 words per workgroup
};

struct shg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int shg_batch_plan_create(shg_context* ctx, const shg_limb_job* host_jobs, int num_jobs,
                          int variant, shg_batch_plan** out) {
  if (!ctx || !out) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (num_jobs < 0 || (num_jobs > 0 && !host_jobs))
    return fail(SHG_E_INVALID_ARGUMENT, "jobs must be a non-negative count with a job array.");
  if (variant < SHG_NTT_AUTO || variant > SHG_NTT_F64)
    return fail(SHG_E_INVALID_ARGUMENT, "unknown transform variant %d.", variant);
  if (capturing(ctx)) return refuse_capture(ctx, "shg_batch_plan_create");
  std::vector<shg::LimbJobDev> jobs(num_jobs);
  bool f64_ok = true, has_other = true;
  int max_log_n = shg::kBatchMinLogN;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (int i = 0; i < num_jobs; ++i) {
      const shg_limb_job& j = host_jobs[i];
      if (j.prime < 0 || j.prime >= (int)ctx->primes.size())
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: invalid prime handle %d.", i, j.prime);
      if (j.log_n < shg::kBatchMinLogN || j.log_n > shg::kBatchMaxLogN)
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: batched jobs need 2^%d <= n <= 2^%d.", i,
                    shg::kBatchMinLogN, shg::kBatchMaxLogN);
      if (ctx->primes[j.prime]->log_n != j.log_n)
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: prime was created for another ring degree.", i);
      if (!aligned16(j.data) || (j.other && !aligned16(j.other)))
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: data must be 16-byte aligned device memory.", i);
      if (ctx->primes[j.prime]->q >= (1ull << 50)) f64_ok = false;
      has_other &= j.other != nullptr;
      max_log_n = j.log_n > max_log_n ? j.log_n : max_log_n;
      jobs[i] = shg::LimbJobDev{j.data, j.other, j.prime, j.log_n};
    }
  }
  if (variant == SHG_NTT_F64 && !f64_ok)
    return fail(SHG_E_INVALID_ARGUMENT, "FP64 transform requires moduli below 2^50.");
  int v = variant;
  if (v == SHG_NTT_AUTO) {
    v = f64_ok ? SHG_NTT_F64 : SHG_NTT_INT_SHOUP;
    auto it = ctx->tuned.find({max_log_n, -1, -1, 0});
    if (f64_ok && it != ctx->tuned.end()) v = it->second & 3;
  }
  // Largest rings first so the static stride balances the workgroups;
  // then by prime for twiddle locality.
  std::stable_sort(jobs.begin(), jobs.end(), [](const shg::LimbJobDev& a, const shg::LimbJobDev& b) {
    return a.log_n != b.log_n ? a.log_n > b.log_n : a.prime < b.prime;
  });
  auto plan = std::make_unique<shg_batch_plan>();
  plan->num_jobs = num_jobs;
  plan->variant = v;
  plan->has_other = has_other;
  plan->shmem = shg::batch_shmem_bytes(max_log_n);
  if (num_jobs > 0) {
    const void* k = shg::batch_kernel_ptr(v, false);
    if (plan->shmem > 65536)
      shg::set_max_dynamic_shmem(k, plan->shmem);
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
        cus <= 0)
      cus = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, shg::kBatchThreads, plan->shmem) !=
            hipSuccess || per_cu <= 0)
      per_cu = 1;
    const long long cap = (long long)cus * per_cu;
    plan->grid = (int)(cap < num_jobs ? cap : num_jobs);
    hipError_t e = hipMalloc(&plan->d_jobs, sizeof(shg::LimbJobDev) * num_jobs);
    if (e != hipSuccess) return fail(SHG_E_NO_MEMORY, "job table allocation failed.");
    e = hipMemcpy(plan->d_jobs, jobs.data(), sizeof(shg::LimbJobDev) * num_jobs,
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(plan->d_jobs);
      return hip_fail(e, "job table upload");
    }
  }
  *out = plan.release();
  return SHG_OK;
}

int shg_batch_mod_reduce_plan_create(shg_context* ctx, const shg_poly_job* host_jobs,
                                     int num_jobs, uint64_t t, int is_ntt, int variant,
                                     shg_batch_plan** out) {
  if (!ctx || !out) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (num_jobs < 0 || (num_jobs > 0 && !host_jobs))
    return fail(SHG_E_INVALID_ARGUMENT, "jobs must be a non-negative count with a job array.");
  if (t == 0) return fail(SHG_E_INVALID_ARGUMENT, "plaintext modulus must be positive.");
  if (variant < SHG_NTT_AUTO || variant > SHG_NTT_F64)
    return fail(SHG_E_INVALID_ARGUMENT, "unknown transform variant %d.", variant);
  if (capturing(ctx)) return refuse_capture(ctx, "shg_batch_mod_reduce_plan_create");
  std::vector<shg::PolyJobDev> jobs(num_jobs);
  bool f64_ok = true;
  int max_log_n = shg::kBatchMinLogN;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (int i = 0; i < num_jobs; ++i) {
      const shg_poly_job& j = host_jobs[i];
      const int L = j.num_limbs;
      if (L <= 1)
        return fail(SHG_E_INVALID_ARGUMENT,
                    "job %d: ModReduceLsb cannot apply with only one prime modulus.", i);
      if (L > shg::kBatchMaxPolyLimbs)
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: at most %d limbs per polynomial.", i,
                    shg::kBatchMaxPolyLimbs);
      if (!j.primes || !j.host_ql_inv)
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: `ql_inv` must be defined with respect to %d moduli.",
                    i, L - 1);
      if (j.log_n < shg::kBatchMinLogN || j.log_n > shg::kBatchMaxLogN)
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: batched jobs need 2^%d <= n <= 2^%d.", i,
                    shg::kBatchMinLogN, shg::kBatchMaxLogN);
      if (!aligned16(j.data))
        return fail(SHG_E_INVALID_ARGUMENT, "job %d: data must be 16-byte aligned device memory.", i);
      shg::PolyJobDev d{};
      d.data = j.data;
      d.log_n = j.log_n;
      d.num_limbs = L;
      d.is_ntt = is_ntt ? 1 : 0;
      for (int l = 0; l < L; ++l) {
        const int p = j.primes[l];
        if (p < 0 || p >= (int)ctx->primes.size())
          return fail(SHG_E_INVALID_ARGUMENT, "job %d: invalid prime handle %d.", i, p);
        if (ctx->primes[p]->log_n != j.log_n)
          return fail(SHG_E_INVALID_ARGUMENT, "job %d: prime was created for another ring degree.", i);
        if (ctx->primes[p]->q >= (1ull << 50)) f64_ok = false;
        d.prime[l] = p;
      }
      // The constants of shg_mod_reduce_lsb (rns/rns_polynomial.cc:262-300).
      const Params* Pl = ctx->primes[d.prime[L - 1]]->mp.get();
      const uint64_t r = Pl->modulus % t;
      const uint64_t k = (Pl->modulus - r) / t;
      d.k_mod_ql = ModInt::ImportInt(k, Pl).value().GetMontgomeryRepresentation();
      for (int l = 0; l < L - 1; ++l) {
        const Params* Pi = ctx->primes[d.prime[l]]->mp.get();
        d.r_mod_qi[l] = ModInt::ImportInt(r, Pi).value().GetMontgomeryRepresentation();
        d.t_mod_qi[l] = ModInt::ImportInt(t, Pi).value().GetMontgomeryRepresentation();
        d.ql_inv[l] = j.host_ql_inv[l];
      }
      if (f64_ok) fill_modreduce_f64(ctx, d, k);
      max_log_n = j.log_n > max_log_n ? j.log_n : max_log_n;
      jobs[i] = d;
    }
  }
  if (variant == SHG_NTT_F64 && !f64_ok)
    return fail(SHG_E_INVALID_ARGUMENT, "FP64 transform requires moduli below 2^50.");
  int v = variant == SHG_NTT_AUTO ? (f64_ok ? SHG_NTT_F64 : SHG_NTT_INT_SHOUP) : variant;
  // Most expensive polynomials (limbs x ring size) first.
  std::stable_sort(jobs.begin(), jobs.end(), [](const shg::PolyJobDev& a, const shg::PolyJobDev& b) {
    return ((long long)a.num_limbs << a.log_n) > ((long long)b.num_limbs << b.log_n);
  });
  auto plan = std::make_unique<shg_batch_plan>();
  plan->modreduce = true;
  plan->num_jobs = num_jobs;
  plan->variant = v;
  plan->has_other = false;
  plan->shmem = shg::batch_shmem_bytes(max_log_n);
  if (num_jobs > 0) {
    const void* k = shg::batch_kernel_ptr(v, true);
    if (plan->shmem > 65536) shg::set_max_dynamic_shmem(k, plan->shmem);
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
        cus <= 0)
      cus = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, shg::kBatchThreads, plan->shmem) !=
            hipSuccess || per_cu <= 0)
      per_cu = 1;
    const long long cap = (long long)cus * per_cu;
    plan->grid = (int)(cap < num_jobs ? cap : num_jobs);
    hipError_t e = hipMalloc(&plan->d_jobs, sizeof(shg::PolyJobDev) * num_jobs);
    if (e != hipSuccess) return fail(SHG_E_NO_MEMORY, "job table allocation failed.");
    e = hipMalloc(&plan->scratch, ((size_t)plan->grid << 16) * 8);
    if (e != hipSuccess) {
      (void)hipFree(plan->d_jobs);
      return fail(SHG_E_NO_MEMORY, "batch scratch allocation failed.");
    }
    e = hipMemcpy(plan->d_jobs, jobs.data(), sizeof(shg::PolyJobDev) * num_jobs,
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(plan->d_jobs);
      (void)hipFree(plan->scratch);
      return hip_fail(e, "job table upload");
    }
  }
  *out = plan.release();
  return SHG_OK;
}

int shg_batch_plan_destroy(shg_context* ctx, shg_batch_plan* plan) {
  if (!ctx || !plan) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_batch_plan_destroy");
  if (plan->d_jobs || plan->scratch) (void)hipStreamSynchronize(ctx->stream);
  if (plan->d_jobs) (void)hipFree(plan->d_jobs);
  if (plan->scratch) (void)hipFree(plan->scratch);
  delete plan;
  return SHG_OK;
}

int shg_batch_run(shg_context* ctx, const shg_batch_plan* plan, int op) {
  if (!ctx || !plan) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (plan->modreduce) {
    if (op != SHG_BATCH_MOD_REDUCE)
      return fail(SHG_E_INVALID_ARGUMENT, "a modulus-switch plan runs SHG_BATCH_MOD_REDUCE only.");
    ProfScope prof(ctx, "batch");
    hipError_t e = shg::launch_batch_modreduce((const shg::PolyJobDev*)plan->d_jobs,
                                               plan->num_jobs, plan->variant, ctx->d_primes,
                                               plan->scratch, plan->grid, plan->shmem, ctx->stream);
    return e == hipSuccess ? SHG_OK : hip_fail(e, "batch launch");
  }
  if (op != SHG_BATCH_NTT_FWD && op != SHG_BATCH_NTT_INV && op != SHG_BATCH_MUL)
    return fail(SHG_E_INVALID_ARGUMENT, "unknown batch op %d.", op);
  if (op == SHG_BATCH_MUL && !plan->has_other)
    return fail(SHG_E_INVALID_ARGUMENT, "multiply jobs need a second operand.");
  ProfScope prof(ctx, "batch");
  hipError_t e = shg::launch_batch_limbs((const shg::LimbJobDev*)plan->d_jobs, plan->num_jobs, op,
                                         plan->variant, ctx->d_primes, plan->grid, plan->shmem,
                                         ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "batch launch");
}

int shg_batch_limbs(shg_context* ctx, const shg_limb_job* host_jobs, int num_jobs, int op) {
  if (ctx && capturing(ctx)) return refuse_capture(ctx, "shg_batch_limbs");
  shg_batch_plan* plan = nullptr;
  int rc = shg_batch_plan_create(ctx, host_jobs, num_jobs, SHG_NTT_AUTO, &plan);
  if (rc) return rc;
  rc = shg_batch_run(ctx, plan, op);
  if (rc == SHG_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(e, "batch");
  }
  const std::string msg = g_last_error;
  shg_batch_plan_destroy(ctx, plan);
  if (rc) g_last_error = msg;
  return rc;
}

int shg_graph_begin(shg_context* ctx) {
  if (!ctx) return fail(SHG_E_INVALID_ARGUMENT, "context must not be null.");
  if (capturing(ctx)) return fail(SHG_E_FAILED_PRECONDITION, "a capture is already in progress.");
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "begin capture");
}

int shg_graph_end(shg_context* ctx, shg_graph** out) {
  if (!ctx || !out) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  auto g = std::make_unique<shg_graph>();
  hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
  if (e != hipSuccess) return hip_fail(e, "end capture");
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    return hip_fail(e, "graph instantiate");
  }
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    ++ctx->live_graphs;
  }
  *out = g.release();
  return SHG_OK;
}

int shg_graph_launch(shg_context* ctx, const shg_graph* graph) {
  if (!ctx || !graph) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  hipError_t e = hipGraphLaunch(graph->exec, ctx->stream);
  return e == hipSuccess ? SHG_OK : hip_fail(e, "graph launch");
}

int shg_graph_destroy(shg_context* ctx, shg_graph* graph) {
  if (!ctx || !graph) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_graph_destroy");
  (void)hipStreamSynchronize(ctx->stream);
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ctx->live_graphs > 0 && --ctx->live_graphs == 0) {
    // No graph can replay a launch that uses a retired buffer any more.
    for (void* p : ctx->retired) (void)hipFree(p);
    ctx->retired.clear();
  }
  return SHG_OK;
}

namespace {

// Candidate configurations for (log n, width class): every kernel family,
// register radix, stage split and prefetch setting that applies at that
// ring degree (flags that a degree ignores are left out).
std::vector<int> ntt_candidates(int log_n, int wide) {
  std::vector<int> bases = {SHG_NTT_INT_SHOUP};
  if (!wide) bases.push_back(SHG_NTT_F64);
  std::vector<int> out;
  for (int b : bases) {
    const bool f64 = b == SHG_NTT_F64;
    out.push_back(b);
    if (log_n >= 12 && log_n <= 14) out.push_back(b | shg::kNttCfgAlt);
    if (f64 && log_n <= 14) {
      out.push_back(b | shg::kNttCfgPrefetch);
      if (log_n >= 12) out.push_back(b | shg::kNttCfgAlt | shg::kNttCfgPrefetch);
    }
    if (f64 && log_n >= 12 && log_n <= 14) out.push_back(b | shg::kNttCfgWave);
    if (log_n == 14) out.push_back(b | shg::kNttCfgSplit);
    if (log_n >= 13 && log_n <= 15) out.push_back(b | shg::kNttCfgHalf);
  }
  return out;
}

// Times `cands` on `limbs` limbs at `buf` (forward + inverse pairs: one
// warm-up, then three interleaved rounds of four pairs; a candidate's time
// is its best round, which discounts clock ramps and other transient noise)
// and returns the fastest.
int time_candidates(shg_context* ctx, const std::vector<int>& cands, uint64_t* buf,
                    const LimbPrimes& lp, int num_limbs, int rows, int log_n, hipError_t* err) {
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipError_t e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  hipStream_t s = ctx->stream;
  auto run = [&](int cfg) {
    hipError_t r = shg::launch_ntt(false, cfg, buf, ctx->d_primes, lp, rows, num_limbs, log_n,
                                   shg::kMaxLogM, s);
    if (r == hipSuccess)
      r = shg::launch_ntt(true, cfg, buf, ctx->d_primes, lp, rows, num_limbs, log_n,
                          shg::kMaxLogM, s);
    return r;
  };
  std::vector<float> best_round(cands.size(), 1e30f);
  for (size_t c = 0; c < cands.size() && e == hipSuccess; ++c) e = run(cands[c]);
  for (int round = 0; round < 3 && e == hipSuccess; ++round) {
    for (size_t c = 0; c < cands.size() && e == hipSuccess; ++c) {
      e = hipEventRecord(ev0, s);
      for (int r = 0; r < 4 && e == hipSuccess; ++r) e = run(cands[c]);
      if (e == hipSuccess) e = hipEventRecord(ev1, s);
      if (e == hipSuccess) e = hipEventSynchronize(ev1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
      if (e == hipSuccess && ms < best_round[c]) best_round[c] = ms;
    }
  }
  int best = cands[0];
  float best_ms = 1e30f;
  for (size_t c = 0; c < cands.size(); ++c) {
    if (best_round[c] < best_ms) {
      best_ms = best_round[c];
      best = cands[c];
    }
  }
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  *err = e;
  return best;
}

uint64_t splitmix(uint64_t& st) {
  st += 0x9e3779b97f4a7c15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

int shg_autotune_ntt(shg_context* ctx, int log_n, int limbs, int* chosen) {
  if (!ctx || !chosen) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  if (log_n < 1 || log_n > 15) return fail(SHG_E_INVALID_ARGUMENT, "log_n must be in [1, 15].");
  if (limbs <= 0) return fail(SHG_E_INVALID_ARGUMENT, "limbs must be positive.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_autotune_ntt");
  // A modulus of the shape the configurations serve: the largest prime
  // below 2^50 with q = 1 mod 2n.
  const uint64_t two_n = 2ull << log_n;
  uint64_t q = ((1ull << 50) - 1) / two_n * two_n + 1;
  while (q > two_n && !is_prime_u64(q)) q -= two_n;
  int pid = -1;
  int rc = shg_prime_create(ctx, q, log_n, &pid);
  if (rc) return rc;
  const size_t n = (size_t)1 << log_n;
  const size_t words = (size_t)limbs * n;
  std::vector<uint64_t> host(words);
  uint64_t st = 0x9e3779b97f4a7c15ull;
  for (auto& w : host) w = splitmix(st) % q;
  void* buf = nullptr;
  if (hipMalloc(&buf, words * 8) != hipSuccess)
    return fail(SHG_E_NO_MEMORY, "autotune buffer of %zu bytes failed.", words * 8);
  hipError_t e = hipMemcpy(buf, host.data(), words * 8, hipMemcpyHostToDevice);
  LimbPrimes lp{};
  lp.id[0] = pid;
  int best = SHG_NTT_F64;
  if (e == hipSuccess)
    best = time_candidates(ctx, ntt_candidates(log_n, 0), (uint64_t*)buf, lp, 1, limbs, log_n, &e);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return hip_fail(e, "autotune");
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int bucket = rows_bucket(limbs);
    ctx->tuned[{log_n, -1, bucket, 0}] = best;
    ctx->tuned[{log_n, -1, -1, 0}] = best;  // batch plans of this ring degree
  }
  *chosen = best;
  return SHG_OK;
}

int shg_autotune_ntt_moduli(shg_context* ctx, const int* primes, int num_limbs, int count,
                            int* chosen) {
  if (!ctx || !chosen) return fail(SHG_E_INVALID_ARGUMENT, "null argument.");
  LimbPrimes lp{};
  int log_n = 0;
  int rc = check_primes(ctx, primes, num_limbs, &log_n, &lp);
  if (rc) return rc;
  if (count <= 0) return fail(SHG_E_INVALID_ARGUMENT, "count must be positive.");
  if (capturing(ctx)) return refuse_capture(ctx, "shg_autotune_ntt_moduli");
  const int wide = width_class(ctx, primes, num_limbs);
  const size_t n = (size_t)1 << log_n;
  const size_t words = (size_t)count * num_limbs * n;
  std::vector<uint64_t> host(words);
  uint64_t st = 0x2545f4914f6cdd1dull;
  for (size_t w = 0; w < words; ++w)
    host[w] = splitmix(st) % ctx->primes[primes[(w / n) % num_limbs]]->q;
  void* buf = nullptr;
  if (hipMalloc(&buf, words * 8) != hipSuccess)
    return fail(SHG_E_NO_MEMORY, "autotune buffer of %zu bytes failed.", words * 8);
  hipError_t e = hipMemcpy(buf, host.data(), words * 8, hipMemcpyHostToDevice);
  const int rows = count * num_limbs;
  int best = wide ? SHG_NTT_INT_SHOUP : SHG_NTT_F64;
  if (e == hipSuccess)
    best = time_candidates(ctx, ntt_candidates(log_n, wide), (uint64_t*)buf, lp, num_limbs, rows,
                           log_n, &e);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return hip_fail(e, "autotune");
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tuned[{log_n, num_limbs, rows_bucket(rows), wide}] = best;
  }
  *chosen = best;
  return SHG_OK;
}

}  // extern "C"
