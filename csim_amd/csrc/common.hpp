// Shared host/device definitions of the csx library (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/csx.h"

namespace csx {

// Thread-local error text behind csx_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define CSX_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::csx::fail(CSX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CSX_TRY(expr)                   \
  do {                                  \
    int rc_ = (expr);                   \
    if (rc_ != CSX_OK) return rc_;      \
  } while (0)

// Device buffer owned by the library (RAII, no copies).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
  // Grows (never shrinks) to at least n bytes; contents are not preserved.
  int reserve(size_t n) {
    if (n <= bytes) return CSX_OK;
    release();
    if (n == 0) return CSX_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) { p = nullptr; return fail(CSX_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
    bytes = n;
    return CSX_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace csx

// The context: one HIP device, one stream, scratch buffers.
struct csx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  csx::DevBuf work;       // operator-API scratch
  csx::DevBuf small;      // small per-call parameter blocks
  csx::DevBuf small2;
};
