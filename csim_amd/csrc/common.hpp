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

// Device memory a caller hands over (a workspace): consecutive 256-byte aligned slices.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, off = 0;
  static size_t round(size_t n) { return (n + 255) / 256 * 256; }
};

// Device buffer owned by the library (RAII, no copies), or a slice of an Arena.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;   // set: reserve() takes slices of it and nothing is freed
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), arena(o.arena) { o.p = nullptr; o.bytes = 0; o.arena = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; bytes = o.bytes; arena = o.arena;
      o.p = nullptr; o.bytes = 0; o.arena = nullptr;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() { if (p && !arena) (void)hipFree(p); p = nullptr; bytes = 0; }
  // Grows (never shrinks) to at least n bytes; contents are not preserved.  With an
  // arena the buffer becomes the arena's next slice (no allocation).
  int reserve(size_t n) {
    if (n <= bytes) return CSX_OK;
    if (arena) {
      const size_t need = Arena::round(n);
      if (arena->off + need > arena->cap)
        return fail(CSX_ERR_NOMEM, "workspace too small: " + std::to_string(arena->off + need) + " bytes needed, " +
                                       std::to_string(arena->cap) + " given");
      p = arena->base + arena->off;
      arena->off += need;
      bytes = n;
      return CSX_OK;
    }
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
