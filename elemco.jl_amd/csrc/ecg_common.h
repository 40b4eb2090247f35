// Shared host/device definitions of the elemco HIP library (gfx950 / MI355X).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace ecg {

// Error raised by every failing HIP call or argument check; the C-ABI layer
// turns it into a negative status code plus a message (ecg_last_error).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define ECG_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::ecg::Error(-3, std::string("HIP error '") + hipGetErrorString(e_) +      \
                                 "' at " __FILE__ ":" + std::to_string(__LINE__));     \
  } while (0)

#define ECG_CHECK(cond, msg)                                                           \
  do {                                                                                 \
    if (!(cond)) throw ::ecg::Error(-2, std::string(msg));                             \
  } while (0)

// Arithmetic used by the tensor contractions.
enum class Prec : int { F64 = 0, F32 = 1, BF16X3 = 2 };

// Device buffer owned by the library (hipMalloc'd, freed on destruction).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b) ECG_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  void ensure(size_t b) { if (b > bytes) alloc(b); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  double* d() const { return static_cast<double*>(p); }
  float* f() const { return static_cast<float*>(p); }
};

}  // namespace ecg
