// common.hpp -- shared helpers of the nosh_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

namespace nosh {

// Error carrying a C-ABI status code; converted at the ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kInvalid = 1,
  kUnsupported = 2,
  kMissingParam = 3,
  kHip = 4,
  kNotBuilt = 5,
  kNotConverged = 6,
};

#define NOSH_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::nosh::Error(::nosh::kHip, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define NOSH_CHECK(cond, code, msg)                 \
  do {                                              \
    if (!(cond)) throw ::nosh::Error((code), (msg)); \
  } while (0)

// Owning device allocation.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) NOSH_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    NOSH_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s) const {
    NOSH_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n) NOSH_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  size_t bytes() const { return n * sizeof(T); }
};

inline unsigned grid_for(int64_t work, int block) {
  int64_t g = (work + block - 1) / block;
  return unsigned(g < 1 ? 1 : g);
}

// Library-wide stream: one per process, default priority.
hipStream_t default_stream();

}  // namespace nosh
