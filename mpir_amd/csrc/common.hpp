// Shared HIP utilities: error handling, device buffers, precision tags.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpir {

// Precision tags used across the C ABI (values are part of the ABI).
enum Prec : int { kF64 = 0, kF32 = 1, kF16 = 2, kBF16 = 3 };

inline const char* prec_name(int p) {
  switch (p) {
    case kF64: return "f64";
    case kF32: return "f32";
    case kF16: return "f16";
    case kBF16: return "bf16";
    default: return "?";
  }
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes returned through the C ABI.
enum Status : int {
  kOk = 0,
  kErrInvalid = -100,   // bad argument / input contract violated
  kErrHip = -101,       // HIP runtime failure
  kErrNoDevice = -102,  // no usable gfx950 device
  kErrState = -103,     // call order (e.g. solve before compute)
  kErrNumeric = -104,   // zero pivot / non-finite values
  kErrInternal = -105,
};

#define MPIR_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::mpir::Error(::mpir::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                               " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define MPIR_CHECK(cond, msg)                                   \
  do {                                                          \
    if (!(cond)) throw ::mpir::Error(::mpir::kErrInvalid, msg); \
  } while (0)

// Owning device allocation (untyped). Size in bytes.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr; o.n_ = o.cap_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; cap_ = o.cap_; o.p_ = nullptr; o.n_ = o.cap_ = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  // Contents are not preserved. An allocation that already holds `bytes`
  // is kept (buffers only grow): hipFree waits for the whole device, so
  // handles that are re-analyzed for every batch of systems reuse their
  // memory instead, and a buffer that has to grow at least doubles, so that
  // growing stays rare (a serving lane's first batches are small).
  void alloc(size_t bytes) {
    if (p_ && bytes <= cap_) {
      n_ = bytes;
      return;
    }
    const size_t want = p_ ? std::max(bytes, 2 * cap_) : bytes;
    if (p_) regrowths().fetch_add(1, std::memory_order_relaxed);
    release();
    if (bytes == 0) return;
    MPIR_HIP(hipMalloc(&p_, want));
    n_ = bytes;
    cap_ = want;
  }
  // device buffers of the process that had to grow (each waits for the device)
  static std::atomic<int64_t>& regrowths() {
    static std::atomic<int64_t> n{0};
    return n;
  }
  // grow-only reallocation (contents not preserved)
  void ensure(size_t bytes) {
    if (bytes > cap_) alloc(bytes);
    else if (bytes > n_) n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = cap_ = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p_); }
  void* get() const { return p_; }
  size_t bytes() const { return n_; }
  size_t capacity() const { return cap_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

template <class T>
inline void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.ensure(v.size() * sizeof(T));
  if (!v.empty()) MPIR_HIP(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

void set_last_error(const std::string& msg);

}  // namespace mpir
