// Device helpers shared by the GEMM and pointwise kernels: global-address-space
// views of generic pointers and the split of complex values into BF16 planes.
#pragma once
#include <hip/hip_runtime.h>

namespace csx {

// FP32 -> BF16, round to nearest even (the hardware conversion, v_cvt_pk_bf16_f32).  The
// empty asm makes x an FP32 value the compiler cannot see through: a narrowing from FP64
// before it stays its own rounding (two roundings, FP64 -> FP32 -> BF16), not folded into
// one FP64 -> BF16 rounding.
__device__ __forceinline__ unsigned short bf16_rne(float x) {
  asm("" : "+v"(x));
  return __builtin_bit_cast(unsigned short, (__bf16)x);
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
// Two-term planes only; three-term planes use store_planes3 below.
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

// ---- three-term planes with a fixed-point leading term --------------------------
// A contraction reads its operands in K windows of kPlaneWindow consecutive values
// (aligned at multiples of kPlaneWindow).  In three-term planes the leading part of
// every value of a window is an integer multiple of one power of two, the window's
// unit u = 2^(e - 8) with 2^(e-1) <= max(|re|, |im|) over the window < 2^e, so
// |hi| <= 256 u.  Products of two leading parts are then integer multiples of
// u_A u_B below 2^16 u_A u_B, and the contraction's sums of them over one window
// (64 products per output) stay below 2^22 u_A u_B: exact in FP32 whatever the
// order of the additions.  The remainder x - hi (|.| <= u / 2) is split into two
// BF16 parts by rounding to nearest, which leaves about 2^-26 of the window's
// maximum.
constexpr int kPlaneWindow = 32;

// e with 2^(e-1) <= m < 2^e (0 for m = 0); m may be rounded up (a coarser unit is safe).
__device__ __forceinline__ int window_exponent(float m) {
  int e = 0;
  (void)frexpf(m, &e);
  return m > 0.0f ? e : 0;
}
// Maximum of m over the LANES consecutive lanes (a power of two, aligned) of a window.
template <int LANES>
__device__ __forceinline__ float window_max(float m) {
#pragma unroll
  for (int o = 1; o < LANES; o <<= 1) m = fmaxf(m, __shfl_xor(m, o, LANES));
  return m;
}
// Three parts of x on the window exponent e: hi = rint(x / u) u, then the FP32 value
// of x - hi (exact in FP64) split into two BF16 parts by round to nearest even.
__device__ __forceinline__ void split3(double x, int e, unsigned short* h) {
  const double hi = ldexp(rint(ldexp(x, 8 - e)), e - 8);
  float r = (float)(x - hi);
  h[0] = bf16_rne((float)hi);      // an integer of at most 9 bits times 2^(e-8): exact
  h[1] = bf16_rne(r);
  r -= bf16_val(h[1]);
  h[2] = bf16_rne(r);
}
template <typename Ptr>
__device__ __forceinline__ void store_planes3(Ptr base, long off, long stride, double vr, double vi, int e) {
  unsigned short hr[3], hi[3];
  split3(vr, e, hr);
  split3(vi, e, hi);
#pragma unroll
  for (int t = 0; t < 3; t++) {
    base[(2 * t) * stride + off] = hr[t];
    base[(2 * t + 1) * stride + off] = hi[t];
  }
}
// Planes of `terms` parts, the window exponent e used by three-term planes only.
template <typename Ptr>
__device__ __forceinline__ void store_planes_w(Ptr base, long off, long stride, int terms, double vr, double vi,
                                               int e) {
  if (terms == 3) store_planes3(base, off, stride, vr, vi, e);
  else store_planes(base, off, stride, terms, vr, vi);
}

}  // namespace csx
