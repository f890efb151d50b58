// Device helpers shared by the GEMM and pointwise kernels: global-address-space
// views of generic pointers and the split of complex values into BF16 planes.
#pragma once
#include <hip/hip_runtime.h>

namespace csx {

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(r >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// Pointers read from the problem array are generic; casting them to the global
// address space makes the loads global_load (vmcnt only) instead of flat_load,
// which also counts against lgkmcnt and so stalls every LDS wait behind HBM.
#define GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ const GLOBAL T* gin(const void* p) { return (const GLOBAL T*)p; }
template <typename T> __device__ __forceinline__ GLOBAL T* gout(void* p) { return (GLOBAL T*)p; }
// builtin vector types for complex values behind global pointers (the HIP vector
// classes have no address-space-qualified copy)
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// (re, im) split into `terms` BF16 parts each: x = hi + mid + lo, every part the
// round-to-nearest-even of what the previous ones leave (from the FP32 value).
template <typename Ptr>
__device__ __forceinline__ void store_planes(Ptr base, long off, long stride, int terms, double vr, double vi) {
  float xr = (float)vr, xi = (float)vi;
  for (int t = 0; t < terms; t++) {
    unsigned short hr = bf16_rne(xr), hi = bf16_rne(xi);
    base[(2 * t) * stride + off] = hr;
    base[(2 * t + 1) * stride + off] = hi;
    xr -= bf16_val(hr);
    xi -= bf16_val(hi);
  }
}

}  // namespace csx
