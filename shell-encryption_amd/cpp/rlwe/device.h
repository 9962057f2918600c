// Process-wide device context and device buffers behind the rlwe C++ API.
// The reference API takes no device argument, so every object uses one
// shg_context per process (device from SHG_DEVICE, default 0), created on
// first use.
#ifndef SHG_RLWE_DEVICE_H_
#define SHG_RLWE_DEVICE_H_

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rlwe/status.h"
#include "shell_gpu.h"

namespace rlwe {
namespace device {

inline shg_context* Context() {
  static std::once_flag once;
  static shg_context* ctx = nullptr;
  std::call_once(once, [] {
    const char* dev = std::getenv("SHG_DEVICE");
    if (shg_create(dev ? std::atoi(dev) : 0, &ctx) != SHG_OK)
      throw std::runtime_error(std::string("libshellgpu: ") + shg_last_error());
  });
  return ctx;
}

// Maps a libshellgpu return code to the reference's Status codes.
inline Status FromCode(int rc) {
  if (rc == SHG_OK) return OkStatus();
  std::string msg = shg_last_error();
  switch (rc) {
    case SHG_E_INVALID_ARGUMENT: return InvalidArgumentError(msg);
    case SHG_E_FAILED_PRECONDITION: return FailedPreconditionError(msg);
    case SHG_E_NO_MEMORY: return ResourceExhaustedError(msg);
    default: return InternalError(msg);
  }
}

// Owning device allocation of `words` uint64 words.
class Buffer {
 public:
  explicit Buffer(size_t words) : words_(words) {
    void* p = nullptr;
    if (shg_alloc(Context(), words * 8, &p) != SHG_OK)
      throw std::runtime_error(std::string("libshellgpu: ") + shg_last_error());
    ptr_ = static_cast<uint64_t*>(p);
  }
  ~Buffer() {
    if (ptr_) shg_free(Context(), ptr_);
  }
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  uint64_t* data() const { return ptr_; }
  size_t words() const { return words_; }
  Status Upload(const uint64_t* host, size_t words, size_t offset = 0) {
    return FromCode(shg_upload(Context(), ptr_ + offset, host, words * 8));
  }
  Status Download(uint64_t* host, size_t words, size_t offset = 0) const {
    return FromCode(shg_download(Context(), host, ptr_ + offset, words * 8));
  }

 private:
  uint64_t* ptr_ = nullptr;
  size_t words_ = 0;
};

// Device-to-device copy on the context's stream.
inline Status CopyWords(uint64_t* dst, const uint64_t* src, size_t words) {
  return FromCode(shg_copy(Context(), dst, src, words * 8));
}

}  // namespace device
}  // namespace rlwe

#endif  // SHG_RLWE_DEVICE_H_
