// Sine and cosine of FP64 phases for the phasor kernels (mirrors, point sources, hex
// masks, the angular-spectrum transfer function).
#pragma once
#include <hip/hip_runtime.h>

namespace csx {

// sin x and cos x for |x| < 2^40.  Cody-Waite reduction by pi/2 = P1 + P2 (P1 the FP64
// nearest pi/2, P2 the FP64 nearest the rest; the part beyond them is below 2^-106 and
// |k| P3 below 1e-20): k = rint(2x/pi), r = (x - k P1) - k P2 with fused multiply-adds
// (the products exact, one rounding each), |r| <= pi/4 (1 + 2^-50).  On that interval
// the Taylor series about 0, sin to r^17 and cos to r^16 (coefficients 1/n!), leave
// truncation errors below 1e-19 and 2e-18; the Horner sums add about one ulp.  The
// quadrant k mod 4 picks and signs the pair.  Compared with numpy's sin and cos the
// results stay within 3e-16 absolute (the reduction's two roundings and the sums').
// Short, branch-free and light on registers, so it can be unrolled over the values a
// thread holds.
__device__ __forceinline__ void sincos_phase(double x, double* s, double* c) {
  const double P1 = 1.5707963267948966;        // 0x1.921fb54442d18p+0
  const double P2 = 6.123233995736766e-17;     // 0x1.1a62633145c07p-54
  const double k = rint(x * 0.6366197723675814);
  double r = __fma_rn(-k, P1, x);
  r = __fma_rn(-k, P2, r);
  const double r2 = r * r;
  // sin r = r + r^3 (-1/3! + r^2/5! - ... + r^14/17!), Horner from the top
  double ps = 1.0 / 355687428096000.0;         // 1/17!
  ps = __fma_rn(ps, r2, -1.0 / 1307674368000.0);
  ps = __fma_rn(ps, r2, 1.0 / 6227020800.0);
  ps = __fma_rn(ps, r2, -1.0 / 39916800.0);
  ps = __fma_rn(ps, r2, 1.0 / 362880.0);
  ps = __fma_rn(ps, r2, -1.0 / 5040.0);
  ps = __fma_rn(ps, r2, 1.0 / 120.0);
  ps = __fma_rn(ps, r2, -1.0 / 6.0);
  const double sr = __fma_rn(r * r2, ps, r);
  // cos r = 1 + r^2 (-1/2! + r^2/4! - ... + r^14/16!)
  double pc = 1.0 / 20922789888000.0;          // 1/16!
  pc = __fma_rn(pc, r2, -1.0 / 87178291200.0);
  pc = __fma_rn(pc, r2, 1.0 / 479001600.0);
  pc = __fma_rn(pc, r2, -1.0 / 3628800.0);
  pc = __fma_rn(pc, r2, 1.0 / 40320.0);
  pc = __fma_rn(pc, r2, -1.0 / 720.0);
  pc = __fma_rn(pc, r2, 1.0 / 24.0);
  pc = __fma_rn(pc, r2, -0.5);
  const double cr = __fma_rn(pc, r2, 1.0);
  const int q = (int)(k - 4.0 * floor(k * 0.25));   // k mod 4, exact for |k| < 2^53
  const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

}  // namespace csx
