// Status / StatusOr / Span value types for the rlwe API.  The reference uses
// absl::Status, absl::StatusOr and absl::Span (shell_encryption/statusor.h,
// status_macros.h); Abseil is not a dependency here, so these small types
// carry the same codes, messages and accessors (ok(), code(), message(),
// value(), operator*, RLWE_ASSIGN_OR_RETURN, RLWE_RETURN_IF_ERROR).
#ifndef SHG_RLWE_STATUS_H_
#define SHG_RLWE_STATUS_H_

#include <cstddef>
#include <optional>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rlwe {

enum class StatusCode {
  kOk = 0,
  kInvalidArgument = 3,
  kFailedPrecondition = 9,
  kInternal = 13,
  kUnavailable = 14,
  kResourceExhausted = 8,
  kUnknown = 2,
};

class Status {
 public:
  Status() : code_(StatusCode::kOk) {}
  Status(StatusCode code, std::string msg) : code_(code), msg_(std::move(msg)) {}
  bool ok() const { return code_ == StatusCode::kOk; }
  StatusCode code() const { return code_; }
  const std::string& message() const { return msg_; }
  std::string ToString() const { return ok() ? "OK" : msg_; }

 private:
  StatusCode code_;
  std::string msg_;
};

inline Status OkStatus() { return Status(); }
inline Status InvalidArgumentError(std::string m) {
  return Status(StatusCode::kInvalidArgument, std::move(m));
}
inline Status FailedPreconditionError(std::string m) {
  return Status(StatusCode::kFailedPrecondition, std::move(m));
}
inline Status InternalError(std::string m) { return Status(StatusCode::kInternal, std::move(m)); }
inline Status UnknownError(std::string m) { return Status(StatusCode::kUnknown, std::move(m)); }
inline Status ResourceExhaustedError(std::string m) {
  return Status(StatusCode::kResourceExhausted, std::move(m));
}
inline bool IsInvalidArgument(const Status& s) { return s.code() == StatusCode::kInvalidArgument; }

template <typename T>
class StatusOr {
 public:
  StatusOr(const Status& s) : status_(s), has_(false) {  // NOLINT
    if (s.ok()) status_ = InternalError("StatusOr constructed from OkStatus");
  }
  StatusOr(T v) : status_(), has_(true), value_(std::in_place, std::move(v)) {}  // NOLINT
  bool ok() const { return has_; }
  const Status& status() const { return status_; }
  T& value() & {
    Check();
    return *value_;
  }
  const T& value() const& {
    Check();
    return *value_;
  }
  T&& value() && {
    Check();
    return std::move(*value_);
  }
  T& operator*() & { return value(); }
  const T& operator*() const& { return value(); }
  T&& operator*() && { return std::move(*this).value(); }
  T* operator->() { return &value(); }
  const T* operator->() const { return &value(); }

 private:
  void Check() const {
    if (!has_) throw std::runtime_error("StatusOr has no value: " + status_.message());
  }
  Status status_;
  bool has_;
  std::optional<T> value_;
};

// Non-owning view of a contiguous sequence (absl::Span).
template <typename T>
class Span {
 public:
  Span() : p_(nullptr), n_(0) {}
  Span(T* p, size_t n) : p_(p), n_(n) {}
  template <typename U>
  Span(const std::vector<U>& v) : p_(v.data()), n_(v.size()) {}  // NOLINT
  template <typename U>
  Span(std::vector<U>& v) : p_(v.data()), n_(v.size()) {}  // NOLINT
  T* data() const { return p_; }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T& operator[](size_t i) const { return p_[i]; }
  T* begin() const { return p_; }
  T* end() const { return p_ + n_; }
  T& back() const { return p_[n_ - 1]; }
  Span first(size_t k) const { return Span(p_, k); }

 private:
  T* p_;
  size_t n_;
};

}  // namespace rlwe

#define SHG_CONCAT_INNER(a, b) a##b
#define SHG_CONCAT(a, b) SHG_CONCAT_INNER(a, b)
#define RLWE_RETURN_IF_ERROR(expr)        \
  do {                                    \
    ::rlwe::Status _shg_st = (expr);      \
    if (!_shg_st.ok()) return _shg_st;    \
  } while (0)
#define RLWE_ASSIGN_OR_RETURN_IMPL(tmp, lhs, expr) \
  auto tmp = (expr);                               \
  if (!tmp.ok()) return tmp.status();              \
  lhs = std::move(tmp).value()
#define RLWE_ASSIGN_OR_RETURN(lhs, expr) \
  RLWE_ASSIGN_OR_RETURN_IMPL(SHG_CONCAT(_shg_so_, __LINE__), lhs, expr)

#endif  // SHG_RLWE_STATUS_H_
