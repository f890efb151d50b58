// Pointwise, FFT, resampling, hex-mask and reduction kernels (gfx950).
// Fields are complex128, column-major (rows = y), batches field-major.
#include "ops.hpp"
#include "phase.hpp"
#include "planes.hpp"

#include <algorithm>
#include <mutex>
#include <map>
#include <cstdlib>

namespace csx {

static inline int grid_for(long n, int block = 256, int cap = 8192) {
  long g = (n + block - 1) / block;
  return (int)std::max<long>(1, std::min<long>(g, cap));
}

// ----------------------------------------------------------------- pointwise --
__global__ void k_mask(double2* __restrict__ E, long npix, long total, const double* __restrict__ m) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    double w = m[i % npix];
    double2 e = E[i];
    // (er + i ei)(w + i w)
    E[i] = make_double2(e.x * w - e.y * w, e.x * w + e.y * w);
  }
}
int k_apply_mask(double2* E, long npix, int nfield, const double* m, hipStream_t s) {
  long total = npix * nfield;
  hipLaunchKernelGGL(k_mask, dim3(grid_for(total)), dim3(256), 0, s, E, npix, total, m);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

__global__ void k_mirror(double2* __restrict__ E, long npix, long total, const double* __restrict__ h,
                         double sign, const double* __restrict__ lambdas) {
#pragma clang fp contract(off)
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / npix;
    // csim: exp((-mirrorSign*2*2*M_PI/lambda) * i * h), evaluated left to right
    double a = -sign * 2 * 2 * M_PI / lambdas[f];
    double ph = a * h[i - f * npix];
    double sn, cs;
    sincos_phase(ph, &sn, &cs);
    double2 e = E[i];
    E[i] = make_double2(e.x * cs - e.y * sn, e.x * sn + e.y * cs);
  }
}
int k_apply_mirror(double2* E, long npix, int nfield, const double* h, double sign,
                   const double* lambdas_dev, hipStream_t s) {
  long total = npix * nfield;
  hipLaunchKernelGGL(k_mirror, dim3(grid_for(total)), dim3(256), 0, s, E, npix, total, h, sign, lambdas_dev);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

__global__ void k_cfactor(double2* __restrict__ E, long npix, long total, const double2* __restrict__ P, int per) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / npix, p = i - f * npix;
    double2 w = P[(f / per) * npix + p];
    double2 e = E[i];
    E[i] = make_double2(e.x * w.x - e.y * w.y, e.x * w.y + e.y * w.x);
  }
}
int k_apply_cfactor(double2* E, long npix, int nfield, const double2* P, int per, hipStream_t s) {
  long total = npix * nfield;
  hipLaunchKernelGGL(k_cfactor, dim3(grid_for(total)), dim3(256), 0, s, E, npix, total, P, per);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// params per field: flux, tipX, tiltY (tip/tilt already times magnification)
__global__ void k_psrc(double2* __restrict__ E, int rows, int cols, int nfield,
                       const double* __restrict__ prm, double ps) {
#pragma clang fp contract(off)
  long npix = (long)rows * cols, total = npix * nfield;
  double Lx = ps * cols, Ly = ps * rows;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / npix, p = i - f * npix;
    int c = (int)(p / rows), r = (int)(p - (long)c * rows);
    double x = -(Lx - ps) / 2.0 + c * ps;
    double y = -(Ly - ps) / 2.0 + r * ps;
    double flux = prm[3 * f], tipX = prm[3 * f + 1], tiltY = prm[3 * f + 2];
    double v = 2 * x * tipX / Lx + 2 * y * tiltY / Ly;
    double ph = 2.0 * M_PI * v;
    double sn, cs;
    sincos_phase(ph, &sn, &cs);
    E[i] = make_double2(flux * cs, flux * sn);
  }
}
int k_point_sources(double2* E, int rows, int cols, int nfield, const double* params_dev,
                    double pixel_scale, hipStream_t s) {
  long total = (long)rows * cols * nfield;
  hipLaunchKernelGGL(k_psrc, dim3(grid_for(total)), dim3(256), 0, s, E, rows, cols, nfield, params_dev, pixel_scale);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// Per-field constants of a pointwise chain (point source, mirror phase factors).
struct PwField {
  bool gen;
  int first;
  double flux, tx, ty, ps, Lx, Ly;
  double mir[kMaxPwOps];            // mirror phase factor per op (mirror.cpp:66-72)
};
__device__ __forceinline__ PwField pw_field(const PwChain& pc, int f) {
  PwField F;
  F.gen = pc.nops > 0 && pc.op[0].kind == PW_SOURCE;
  F.first = F.gen ? 1 : 0;
  F.flux = F.tx = F.ty = 0;
  if (F.gen) { const double* prm = pc.src_params + 3 * f; F.flux = prm[0]; F.tx = prm[1]; F.ty = prm[2]; }
  F.ps = pc.pixel_scale; F.Lx = F.ps * pc.cols; F.Ly = F.ps * pc.rows;
#pragma unroll
  for (int k = 0; k < kMaxPwOps; k++)
    F.mir[k] = (k < pc.nops && pc.op[k].kind == PW_MIRROR) ? -pc.op[k].sign * 2 * 2 * M_PI / pc.lambdas[f] : 0.0;
  return F;
}
// Value of pixel p = c rows + r (column-major) of field f after the chain: generated
// from the point source, or the value (er, ei) passed in, times every element factor.
// Callers that hold a column fixed pass the same c for all its pixels (the x term is
// then formed once).
__device__ __forceinline__ void pw_point_rc(const PwChain& pc, const PwField& F, int f, int p, int r, int c,
                                            double& er, double& ei) {
#pragma clang fp contract(off)
  if (F.gen) {
    // efield::init_point_sources (src/data/efield.cpp:179-206)
    double x = -(F.Lx - F.ps) / 2.0 + c * F.ps, y = -(F.Ly - F.ps) / 2.0 + r * F.ps;
    double v = 2 * x * F.tx / F.Lx + 2 * y * F.ty / F.Ly;
    double sn, cs;
    sincos_phase(2.0 * M_PI * v, &sn, &cs);
    er = F.flux * cs; ei = F.flux * sn;
  }
  // unrolled over the largest op count: every index constant (no private arrays in memory)
#pragma unroll
  for (int k = 0; k < kMaxPwOps; k++) {
    if (k < F.first || k >= pc.nops) continue;
    const PwOp& o = pc.op[k];
    if (o.kind == PW_MASK_1PI) {           // mask.cpp: E *= (1+i) m
      double w = gin<double>(o.p)[p];
      double tr = er * w - ei * w;
      ei = er * w + ei * w;
      er = tr;
    } else if (o.kind == PW_MIRROR) {
      double ph = F.mir[k] * gin<double>(o.p)[p];
      double sn, cs;
      sincos_phase(ph, &sn, &cs);
      double tr = er * cs - ei * sn;
      ei = er * sn + ei * cs;
      er = tr;
    } else if (o.kind == PW_CFACTOR) {
      f64x2 w = gin<f64x2>(o.p)[(long)(f / o.per) * ((long)pc.rows * pc.cols) + p];
      double tr = er * w.x - ei * w.y;
      ei = er * w.y + ei * w.x;
      er = tr;
    }
  }
}

// pw_point_rc on V pixels of one field at once (pixel m = p[m] at row r[m] of column c),
// element by element: the V map values of an element are loaded before any is used,
// so their loads are in flight together (one wait per element, where a per-pixel chain
// waits once per pixel and element).  The element loop is not unrolled (the code of
// each element kind appears once).  Same arithmetic as pw_point_rc, value for value.
// The first NP (<= 3, <= NX) elements after the source come with their map values
// already loaded (masks and mirrors): element F.first + j takes pm[j][m].
template <int V, int NP, int NX>
__device__ __forceinline__ void pw_points(const PwChain& pc, const PwField& F, int f, const int (&p)[V],
                                          const int (&r)[V], int c, double (&er)[V], double (&ei)[V],
                                          const double (&pm)[NX][V]) {
#pragma clang fp contract(off)
  static_assert(NP <= NX && NP <= 3, "preloaded maps");
  if (F.gen) {
    if (F.tx == 0.0 && F.ty == 0.0) {
      // on-axis source: the phase 2 pi v is +0 at every pixel, sincos_phase(+0) = (+0, 1)
#pragma unroll
      for (int m = 0; m < V; m++) { er[m] = F.flux * 1.0; ei[m] = F.flux * 0.0; }
    } else {
#pragma unroll
      for (int m = 0; m < V; m++) {
        double x = -(F.Lx - F.ps) / 2.0 + c * F.ps, y = -(F.Ly - F.ps) / 2.0 + r[m] * F.ps;
        double v = 2 * x * F.tx / F.Lx + 2 * y * F.ty / F.Ly;
        double sn, cs;
        sincos_phase(2.0 * M_PI * v, &sn, &cs);
        er[m] = F.flux * cs; ei[m] = F.flux * sn;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // E *= (1 + i) w (mask.cpp)
  auto mask = [&](const double (&w)[V]) {
#pragma unroll
    for (int m = 0; m < V; m++) {
      double tr = er[m] * w[m] - ei[m] * w[m];
      ei[m] = er[m] * w[m] + ei[m] * w[m];
      er[m] = tr;
    }
  };
  // E *= exp(i mir h) (mirror.cpp:66-72)
  auto mirror = [&](int k, const double (&h)[V]) {
    const double mir = -pc.op[k].sign * 2 * 2 * M_PI / pc.lambdas[f];   // as pw_field
#pragma unroll
    for (int m = 0; m < V; m++) {
      double sn, cs;
      sincos_phase(mir * h[m], &sn, &cs);
      double tr = er[m] * cs - ei[m] * sn;
      ei[m] = er[m] * sn + ei[m] * cs;
      er[m] = tr;
      // one phasor at a time: V interleaved phasors would not fit the registers
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // the elements whose maps came loaded (masks or mirrors)
#pragma unroll
  for (int q = 0; q < NP; q++) {
    const int k = F.first + q;
    if (pc.op[k].kind == PW_MASK_1PI) mask(pm[q]);
    else mirror(k, pm[q]);
  }
#pragma unroll 1
  for (int k = F.first + NP; k < pc.nops; k++) {
    const int kind = pc.op[k].kind;
    if (kind == PW_MASK_1PI) {
      const GLOBAL double* mp = gin<double>(pc.op[k].p);
      double w[V];
#pragma unroll
      for (int m = 0; m < V; m++) w[m] = mp[p[m]];
      mask(w);
    } else if (kind == PW_MIRROR) {
      const GLOBAL double* hp = gin<double>(pc.op[k].p);
      double h[V];
#pragma unroll
      for (int m = 0; m < V; m++) h[m] = hp[p[m]];
      mirror(k, h);
    } else if (kind == PW_CFACTOR) {
      const GLOBAL f64x2* wp = gin<f64x2>(pc.op[k].p) + (long)(f / pc.op[k].per) * ((long)pc.rows * pc.cols);
      f64x2 w[V];
#pragma unroll
      for (int m = 0; m < V; m++) w[m] = wp[p[m]];
#pragma unroll
      for (int m = 0; m < V; m++) {
        double tr = er[m] * w[m].x - ei[m] * w[m].y;
        ei[m] = er[m] * w[m].y + ei[m] * w[m].x;
        er[m] = tr;
      }
    }
  }
}

__device__ __forceinline__ void pw_point(const PwChain& pc, const PwField& F, int f, int p, double& er, double& ei) {
  const int c = p / pc.rows;
  pw_point_rc(pc, F, f, p, p - c * pc.rows, c, er, ei);
}

// One pass: point-source generation and/or element factors, one field per
// blockIdx.y (per-field constants hoisted), output in place (complex128) or at
// the working precision of a following MFT (complex64 or BF16 planes).
__global__ void k_pointwise(double2* __restrict__ E, PwChain pc) {
#pragma clang fp contract(off)
  const int f = blockIdx.y;
  const long npix = (long)pc.rows * pc.cols;
  const PwField F = pw_field(pc, f);
  GLOBAL f64x2* Ef = gout<f64x2>(E) + (long)f * npix;
  const int np32 = (int)npix;                // fields are far below 2^31 pixels
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np32; p += gridDim.x * blockDim.x) {
    double er = 0.0, ei = 0.0;
    if (!F.gen) {
      f64x2 e = Ef[p];
      er = e.x; ei = e.y;
    }
    pw_point(pc, F, f, p, er, ei);
    if (pc.out_kind == PW_OUT_C128) Ef[p] = f64x2{er, ei};
    else if (pc.out_kind == PW_OUT_C64) gout<f32x2>(pc.out)[(long)f * npix + p] = f32x2{(float)er, (float)ei};
    else {
      // three-term planes: windows of 32 consecutive pixels of a column (rows and the
      // grid stride are multiples of 32, so a window's lanes run the loop together)
      int e = 0;
      if (pc.terms == 3) e = window_exponent(window_max<kPlaneWindow>((float)fmax(fabs(er), fabs(ei))));
      store_planes_w(gout<unsigned short>(pc.out) + (long)f * 2 * pc.terms * npix, p, npix, pc.terms, er, ei, e);
    }
  }
}
// Pixel-major generation pass: a thread owns 2 consecutive pixels of every field.
// The factors that do not depend on the field are formed once (the (1+i) m masks),
// the mirror phasors once per wavelength (the mirror phases summed: one sincos per
// pixel and wavelength however many mirrors), the source phasor once per source;
// each field value is then source x factor(lambda), written at the output precision.
// Differs from applying the elements in sequence only by FP64 rounding.
// Phasor of an FP64 phase.  SPLIT (output in two BF16 planes, 16 significant bits): the
// phase is reduced to [-pi, pi] in FP64 and the sine and cosine taken in FP32, whose
// error (~1e-7) is far below the planes' rounding (2^-17); otherwise sincos_phase.
template <bool SPLIT>
__device__ __forceinline__ void phasor(double ph, double* sn, double* cs) {
  if (SPLIT) {
    const double k = rint(ph * (0.5 / M_PI));
    const float r = (float)(ph - k * (2.0 * M_PI));
    float s, c;
    sincosf(r, &s, &c);
    *sn = s;
    *cs = c;
  } else {
    sincos_phase(ph, sn, cs);
  }
}

template <int NL, bool SPLIT>
__global__ __launch_bounds__(256) void k_pointwise_gen(double2* __restrict__ E, PwChain pc, int nouter) {
#pragma clang fp contract(off)
  constexpr int nl = NL;
  const int npix = pc.rows * pc.cols;
  const int p0 = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (p0 >= npix) return;
  double cr[2] = {1.0, 1.0}, ci[2] = {0.0, 0.0};
  for (int k = 1; k < pc.nops; k++) {
    if (pc.op[k].kind != PW_MASK_1PI) continue;
    const f64x2 w = *reinterpret_cast<const GLOBAL f64x2*>(gin<double>(pc.op[k].p) + p0);
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const double tr = cr[x] * w[x] - ci[x] * w[x];
      ci[x] = cr[x] * w[x] + ci[x] * w[x];
      cr[x] = tr;
    }
  }
  double Fr[NL][2], Fi[NL][2];
#pragma unroll
  for (int li = 0; li < NL; li++) {
    double ph[2] = {0.0, 0.0};
    bool any = false;
    for (int k = 1; k < pc.nops; k++) {
      if (pc.op[k].kind != PW_MIRROR) continue;
      // mirror.cpp:66-72: exp((-mirrorSign*2*2*M_PI/lambda) * i * h)
      const double a = -pc.op[k].sign * 2 * 2 * M_PI / pc.lambdas[li];
      const f64x2 h = *reinterpret_cast<const GLOBAL f64x2*>(gin<double>(pc.op[k].p) + p0);
      ph[0] += a * h[0];
      ph[1] += a * h[1];
      any = true;
    }
#pragma unroll
    for (int x = 0; x < 2; x++) {
      if (any) {
        double sn, cs;
        phasor<SPLIT>(ph[x], &sn, &cs);
        Fr[li][x] = cr[x] * cs - ci[x] * sn;
        Fi[li][x] = cr[x] * sn + ci[x] * cs;
      } else {
        Fr[li][x] = cr[x];
        Fi[li][x] = ci[x];
      }
    }
  }
  const double ps = pc.pixel_scale, Lx = ps * pc.cols, Ly = ps * pc.rows;
  double sr[2] = {0, 0}, si[2] = {0, 0};
  double lflux = 0, ltx = 0, lty = 0;
  for (int o = 0; o < nouter; o++) {
    const double* prm = pc.src_params + 3 * o * nl;
    const double flux = prm[0], tx = prm[1], ty = prm[2];
    if (o == 0 || flux != lflux || tx != ltx || ty != lty) {
      // efield::init_point_sources (src/data/efield.cpp:179-206)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const int p = p0 + x, c = p / pc.rows, r = p - c * pc.rows;
        const double xx = -(Lx - ps) / 2.0 + c * ps, yy = -(Ly - ps) / 2.0 + r * ps;
        const double v = 2 * xx * tx / Lx + 2 * yy * ty / Ly;
        double sn, cs;
        phasor<SPLIT>(2.0 * M_PI * v, &sn, &cs);
        sr[x] = flux * cs;
        si[x] = flux * sn;
      }
      lflux = flux; ltx = tx; lty = ty;
    }
#pragma unroll
    for (int li = 0; li < NL; li++) {
      const long f = (long)o * nl + li;
      double er[2], ei[2];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        er[x] = sr[x] * Fr[li][x] - si[x] * Fi[li][x];
        ei[x] = sr[x] * Fi[li][x] + si[x] * Fr[li][x];
      }
      if (pc.out_kind == PW_OUT_C128) {
        GLOBAL f64x2* d = gout<f64x2>(E) + f * npix + p0;
        d[0] = f64x2{er[0], ei[0]};
        d[1] = f64x2{er[1], ei[1]};
      } else if (pc.out_kind == PW_OUT_C64) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        *reinterpret_cast<GLOBAL f32x4*>(gout<float>(pc.out) + 2 * (f * npix + p0)) =
            f32x4{(float)er[0], (float)ei[0], (float)er[1], (float)ei[1]};
      } else {
        // BF16 planes of field f: plane q at (2 terms f + q) npix; two pixels = one 32-bit store
        GLOBAL unsigned* d = reinterpret_cast<GLOBAL unsigned*>(gout<unsigned short>(pc.out) + f * 2 * pc.terms * npix + p0);
        const int stride = npix / 2;
        if (pc.terms == 3) {
          // windows of 32 pixels = 16 consecutive lanes (rows are a multiple of 32)
          const float mx = (float)fmax(fmax(fabs(er[0]), fabs(ei[0])), fmax(fabs(er[1]), fabs(ei[1])));
          const int e = window_exponent(window_max<kPlaneWindow / 2>(mx));
          unsigned short h[2][2][3];
#pragma unroll
          for (int x = 0; x < 2; x++) { split3(er[x], e, h[x][0]); split3(ei[x], e, h[x][1]); }
#pragma unroll
          for (int t = 0; t < 3; t++) {
            d[(2 * t) * stride] = (unsigned)h[0][0][t] | ((unsigned)h[1][0][t] << 16);
            d[(2 * t + 1) * stride] = (unsigned)h[0][1][t] | ((unsigned)h[1][1][t] << 16);
          }
          continue;
        }
        float xr[2] = {(float)er[0], (float)er[1]}, xi[2] = {(float)ei[0], (float)ei[1]};
        for (int t = 0; t < pc.terms; t++) {
          unsigned short hr[2], hi[2];
#pragma unroll
          for (int x = 0; x < 2; x++) {
            hr[x] = bf16_rne(xr[x]);
            hi[x] = bf16_rne(xi[x]);
            xr[x] -= bf16_val(hr[x]);
            xi[x] -= bf16_val(hi[x]);
          }
          d[(2 * t) * stride] = (unsigned)hr[0] | ((unsigned)hr[1] << 16);
          d[(2 * t + 1) * stride] = (unsigned)hi[0] | ((unsigned)hi[1] << 16);
        }
      }
    }
  }
}

// The generation pass into two-term BF16 planes, 8 pixels per thread: past the FP64
// phases everything is FP32 (its 24 bits are well above the planes' 16).  The
// wavelength loop is outermost: one mirror phasor per pixel and wavelength, the
// source phasors (at most kBf2Outer source / candidate rows) kept in registers,
// and each plane of a field written with one 16-byte store per thread.
constexpr int kBf2Outer = 4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <int PX>
__global__ __launch_bounds__(256) void k_pointwise_gen_bf2(PwChain pc, int nouter) {
#pragma clang fp contract(off)
  static_assert(PX == 4 || PX == 8, "pixels per thread");
  const int nl = pc.nl;
  const int npix = pc.rows * pc.cols;
  const int p0 = PX * (blockIdx.x * blockDim.x + threadIdx.x);
  if (p0 >= npix) return;
  float cr[PX], ci[PX];
#pragma unroll
  for (int x = 0; x < PX; x++) { cr[x] = 1.0f; ci[x] = 0.0f; }
  for (int k = 1; k < pc.nops; k++) {
    if (pc.op[k].kind != PW_MASK_1PI) continue;
    const GLOBAL f64x2* wp = reinterpret_cast<const GLOBAL f64x2*>(gin<double>(pc.op[k].p) + p0);
#pragma unroll
    for (int x2 = 0; x2 < PX / 2; x2++) {      // mask.cpp: E *= (1+i) m
      const f64x2 w2 = wp[x2];
#pragma unroll
      for (int y = 0; y < 2; y++) {
        const int x = 2 * x2 + y;
        const float w = (float)w2[y];
        const float tr = cr[x] * w - ci[x] * w;
        ci[x] = cr[x] * w + ci[x] * w;
        cr[x] = tr;
      }
    }
  }
  // source phasors of every outer row (efield::init_point_sources, src/data/efield.cpp:179-206)
  const double ps = pc.pixel_scale, Lx = ps * pc.cols, Ly = ps * pc.rows;
  float sr[kBf2Outer][PX], si[kBf2Outer][PX];
#pragma unroll
  for (int o = 0; o < kBf2Outer; o++) {
    if (o >= nouter) break;
    const double* prm = pc.src_params + 3 * o * nl;
    const double flux = prm[0], tx = prm[1], ty = prm[2];
#pragma unroll
    for (int x = 0; x < PX; x++) {
      const int p = p0 + x, c = p / pc.rows, r = p - c * pc.rows;
      const double xx = -(Lx - ps) / 2.0 + c * ps, yy = -(Ly - ps) / 2.0 + r * ps;
      const double v = 2 * xx * tx / Lx + 2 * yy * ty / Ly;
      double sn, cs;
      phasor<true>(2.0 * M_PI * v, &sn, &cs);
      sr[o][x] = (float)(flux * cs);
      si[o][x] = (float)(flux * sn);
    }
  }
  const int stride = npix / PX;                // one plane in units of PX bf16 values
  for (int li = 0; li < nl; li++) {
    float Fr[PX], Fi[PX];
    double ph[PX];
#pragma unroll
    for (int x = 0; x < PX; x++) ph[x] = 0.0;
    bool any = false;
    for (int k = 1; k < pc.nops; k++) {
      if (pc.op[k].kind != PW_MIRROR) continue;
      const double a = -pc.op[k].sign * 2 * 2 * M_PI / pc.lambdas[li];   // mirror.cpp:66-72
      const GLOBAL f64x2* hp = reinterpret_cast<const GLOBAL f64x2*>(gin<double>(pc.op[k].p) + p0);
#pragma unroll
      for (int x2 = 0; x2 < PX / 2; x2++) {
        const f64x2 h2 = hp[x2];
        ph[2 * x2] += a * h2[0];
        ph[2 * x2 + 1] += a * h2[1];
      }
      any = true;
    }
#pragma unroll
    for (int x = 0; x < PX; x++) {
      if (any) {
        double sn, cs;
        phasor<true>(ph[x], &sn, &cs);
        Fr[x] = cr[x] * (float)cs - ci[x] * (float)sn;
        Fi[x] = cr[x] * (float)sn + ci[x] * (float)cs;
      } else {
        Fr[x] = cr[x];
        Fi[x] = ci[x];
      }
    }
#pragma unroll
    for (int o = 0; o < kBf2Outer; o++) {
      if (o >= nouter) break;
      const long f = (long)o * nl + li;
      unsigned short h[4][PX];                 // planes: hi re, hi im, lo re, lo im
#pragma unroll
      for (int x = 0; x < PX; x++) {
        const float er = sr[o][x] * Fr[x] - si[o][x] * Fi[x];
        const float ei = sr[o][x] * Fi[x] + si[o][x] * Fr[x];
        h[0][x] = bf16_rne(er);
        h[1][x] = bf16_rne(ei);
        h[2][x] = bf16_rne(er - bf16_val(h[0][x]));
        h[3][x] = bf16_rne(ei - bf16_val(h[1][x]));
      }
      GLOBAL unsigned short* base = gout<unsigned short>(pc.out) + f * 4 * npix + p0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        unsigned w[PX / 2];
#pragma unroll
        for (int x2 = 0; x2 < PX / 2; x2++) w[x2] = (unsigned)h[q][2 * x2] | ((unsigned)h[q][2 * x2 + 1] << 16);
        if constexpr (PX == 8)
          reinterpret_cast<GLOBAL u32x4*>(base)[q * stride] = u32x4{w[0], w[1], w[2], w[3]};
        else
          reinterpret_cast<GLOBAL u32x2*>(base)[q * stride] = u32x2{w[0], w[1]};
      }
    }
  }
}

// The generation pass into three-term planes (the bf16x6 working precision), PX = 4
// pixels per thread: FP64 throughout as k_pointwise_gen (masks, the mirror phases
// summed into one phasor per pixel and wavelength, FP64 sincos), each 32-pixel window's
// exponent agreed over 8 lanes, one 8-byte store per plane.  Wavelengths outer, the
// (at most kBf3Outer) source rows inner with their phasors in registers.
constexpr int kBf3Outer = 2;
template <int PX>
__global__ __launch_bounds__(256) void k_pointwise_gen_bf3(PwChain pc, int nouter) {
#pragma clang fp contract(off)
  static_assert(PX == 4, "pixels per thread: a 32-pixel window is 8 lanes");
  const int nl = pc.nl;
  const int npix = pc.rows * pc.cols;
  const int p0 = PX * (blockIdx.x * blockDim.x + threadIdx.x);
  if (p0 >= npix) return;                      // whole windows: npix is a multiple of 32
  double cr[PX], ci[PX];
#pragma unroll
  for (int x = 0; x < PX; x++) { cr[x] = 1.0; ci[x] = 0.0; }
  for (int k = 1; k < pc.nops; k++) {
    if (pc.op[k].kind != PW_MASK_1PI) continue;
    const GLOBAL f64x2* wp = reinterpret_cast<const GLOBAL f64x2*>(gin<double>(pc.op[k].p) + p0);
#pragma unroll
    for (int x2 = 0; x2 < PX / 2; x2++) {      // mask.cpp: E *= (1+i) m
      const f64x2 w2 = wp[x2];
#pragma unroll
      for (int y = 0; y < 2; y++) {
        const int x = 2 * x2 + y;
        const double w = w2[y];
        const double tr = cr[x] * w - ci[x] * w;
        ci[x] = cr[x] * w + ci[x] * w;
        cr[x] = tr;
      }
    }
  }
  // source phasors of every outer row (efield::init_point_sources, src/data/efield.cpp:179-206)
  const double ps = pc.pixel_scale, Lx = ps * pc.cols, Ly = ps * pc.rows;
  double sr[kBf3Outer][PX], si[kBf3Outer][PX];
#pragma unroll
  for (int o = 0; o < kBf3Outer; o++) {
    if (o >= nouter) break;
    const double* prm = pc.src_params + 3 * o * nl;
    const double flux = prm[0], tx = prm[1], ty = prm[2];
#pragma unroll
    for (int x = 0; x < PX; x++) {
      const int p = p0 + x, c = p / pc.rows, r = p - c * pc.rows;
      const double xx = -(Lx - ps) / 2.0 + c * ps, yy = -(Ly - ps) / 2.0 + r * ps;
      const double v = 2 * xx * tx / Lx + 2 * yy * ty / Ly;
      double sn, cs;
      sincos_phase(2.0 * M_PI * v, &sn, &cs);
      sr[o][x] = flux * cs;
      si[o][x] = flux * sn;
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  const int stride = npix / PX;                // one plane in units of PX bf16 values
  for (int li = 0; li < nl; li++) {
    double Fr[PX], Fi[PX], ph[PX];
#pragma unroll
    for (int x = 0; x < PX; x++) ph[x] = 0.0;
    bool any = false;
    for (int k = 1; k < pc.nops; k++) {
      if (pc.op[k].kind != PW_MIRROR) continue;
      const double a = -pc.op[k].sign * 2 * 2 * M_PI / pc.lambdas[li];   // mirror.cpp:66-72
      const GLOBAL f64x2* hp = reinterpret_cast<const GLOBAL f64x2*>(gin<double>(pc.op[k].p) + p0);
#pragma unroll
      for (int x2 = 0; x2 < PX / 2; x2++) {
        const f64x2 h2 = hp[x2];
        ph[2 * x2] += a * h2[0];
        ph[2 * x2 + 1] += a * h2[1];
      }
      any = true;
    }
#pragma unroll
    for (int x = 0; x < PX; x++) {
      if (any) {
        double sn, cs;
        sincos_phase(ph[x], &sn, &cs);
        Fr[x] = cr[x] * cs - ci[x] * sn;
        Fi[x] = cr[x] * sn + ci[x] * cs;
      } else {
        Fr[x] = cr[x];
        Fi[x] = ci[x];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int o = 0; o < kBf3Outer; o++) {
      if (o >= nouter) break;
      const long f = (long)o * nl + li;
      double er[PX], ei[PX];
      float mx = 0.0f;
#pragma unroll
      for (int x = 0; x < PX; x++) {
        er[x] = sr[o][x] * Fr[x] - si[o][x] * Fi[x];
        ei[x] = sr[o][x] * Fi[x] + si[o][x] * Fr[x];
        mx = fmaxf(mx, (float)fmax(fabs(er[x]), fabs(ei[x])));
      }
      const int e = window_exponent(window_max<kPlaneWindow / PX>(mx));
      unsigned short h[PX][2][3];
#pragma unroll
      for (int x = 0; x < PX; x++) { split3(er[x], e, h[x][0]); split3(ei[x], e, h[x][1]); }
      GLOBAL u32x2* base = reinterpret_cast<GLOBAL u32x2*>(gout<unsigned short>(pc.out) + f * 6 * (long)npix + p0);
#pragma unroll
      for (int t = 0; t < 3; t++)
#pragma unroll
        for (int c = 0; c < 2; c++)
          base[(2 * t + c) * stride] = u32x2{(unsigned)h[0][c][t] | ((unsigned)h[1][c][t] << 16),
                                             (unsigned)h[2][c][t] | ((unsigned)h[3][c][t] << 16)};
    }
  }
}

int pointwise_chain(double2* E, int nfield, const PwChain& pc, hipStream_t s) {
  if (pc.nops <= 0 || nfield <= 0) return CSX_OK;
  if (pc.out_kind != PW_OUT_C128 && !pc.out) return fail(CSX_ERR_ARG, "pointwise: no output buffer");
  long npix = (long)pc.rows * pc.cols;
  if (pc.out_kind == PW_OUT_PLANES && pc.terms == 3 && pc.rows % kPlaneWindow)
    return fail(CSX_ERR_UNSUPPORTED, "pointwise: three-term planes need a row count divisible by 32");
  if (pc.nl > 0) {
    if (pc.op[0].kind != PW_SOURCE || pc.nl > kMaxPwLam || npix % 2 || nfield % pc.nl)
      return fail(CSX_ERR_ARG, "pointwise: pixel-major pass preconditions");
    for (int k = 1; k < pc.nops; k++)
      if (pc.op[k].kind != PW_MASK_1PI && pc.op[k].kind != PW_MIRROR)
        return fail(CSX_ERR_ARG, "pointwise: pixel-major pass takes masks and mirrors only");
    typedef void (*GenKernel)(double2*, PwChain, int);
#define CSX_GEN_ROW(S)                                                                                    \
  {k_pointwise_gen<1, S>, k_pointwise_gen<2, S>, k_pointwise_gen<3, S>, k_pointwise_gen<4, S>,            \
   k_pointwise_gen<5, S>, k_pointwise_gen<6, S>, k_pointwise_gen<7, S>, k_pointwise_gen<8, S>,            \
   k_pointwise_gen<9, S>, k_pointwise_gen<10, S>, k_pointwise_gen<11, S>, k_pointwise_gen<12, S>,         \
   k_pointwise_gen<13, S>, k_pointwise_gen<14, S>, k_pointwise_gen<15, S>, k_pointwise_gen<16, S>}
    static const GenKernel kGen[2][kMaxPwLam] = {CSX_GEN_ROW(false), CSX_GEN_ROW(true)};
#undef CSX_GEN_ROW
    static_assert(kMaxPwLam == 16, "one instance per wavelength count");
    if (pc.out_kind == PW_OUT_PLANES && pc.terms == 3 && npix % kPlaneWindow == 0 && nfield / pc.nl <= kBf3Outer) {
      hipLaunchKernelGGL(k_pointwise_gen_bf3<4>, dim3((int)((npix / 4 + 255) / 256)), dim3(256), 0, s, pc,
                         nfield / pc.nl);
      CSX_HIP(hipGetLastError());
      return CSX_OK;
    }
    const int split = pc.out_kind == PW_OUT_PLANES && pc.terms == 2 ? 1 : 0;   // 16-bit planes only
    if (split && npix % 8 == 0 && nfield / pc.nl <= kBf2Outer) {
      // one field per wavelength (sources folded into the next MFT): 4 pixels per thread
      // keep twice the threads in flight (27 us against 31 us for 9 x 1024^2 fields)
      if (nfield == pc.nl)
        hipLaunchKernelGGL(k_pointwise_gen_bf2<4>, dim3((int)((npix / 4 + 255) / 256)), dim3(256), 0, s, pc,
                           nfield / pc.nl);
      else
        hipLaunchKernelGGL(k_pointwise_gen_bf2<8>, dim3((int)((npix / 8 + 255) / 256)), dim3(256), 0, s, pc,
                           nfield / pc.nl);
      CSX_HIP(hipGetLastError());
      return CSX_OK;
    }
    const int nb = (int)((npix / 2 + 255) / 256);
    hipLaunchKernelGGL(kGen[split][pc.nl - 1], dim3(nb), dim3(256), 0, s, E, pc, nfield / pc.nl);
    CSX_HIP(hipGetLastError());
    return CSX_OK;
  }
  int bx = (int)std::min<long>(std::max<long>(1, 4096 / nfield), (npix + 255) / 256);
  hipLaunchKernelGGL(k_pointwise, dim3(bx, nfield), dim3(256), 0, s, E, pc);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ---------------------------------------------------------------------- FFT --
// Stockham autosort transforms of length N (a power of two) by T = max(1, N / 8)
// threads each, the data in registers: between stages thread t holds the V = N / T
// values t + m T (m < V).  Radix 8 while the length allows, then one radix-4 or
// radix-2 stage.  Stage of radix R on sub-transforms of length L (L = 1, R, R R', ...):
//   butterfly b = t + q T (q < V / R), k = b mod L, j = b div L:
//   x_s = v[q + s V / R] w^(s k) = in[b + s N / R] w^(s k), w = exp(sign i 2 pi / (L R));
//   y = DFT_R(x);  out[j L R + k + u L] = y_u.
// Between stages the outputs go through LDS: written at their out index, read back in
// the t + m T layout.  The last stage (L R = N, j = 0) leaves out[b + u N / R] =
// out[t + (q + u V / R) T]: the register layout itself, so a transform reads its input
// from and writes its output to global memory with no LDS pass, and the Fresnel
// step's forward and inverse column transforms meet in registers.
// LDS indices carry one padding slot per 8 values (fft_lds_ix): the first stages'
// stride-8 and stride-64 writes then spread over all banks.
// tw[q] = exp(i 2 pi q / N), q < N (conjugated for sign -1).
constexpr int kFftBlock = 256;           // threads of a row-pass block (several columns below N = 2048)
// A column-pass block loads its panel, transforms it and stores it: two blocks per CU
// (half the LDS and of the 2048 threads each) let one block's loads and stores overlap
// the other's transforms.
constexpr int kFftPanelMaxThreads = 512;
constexpr int kFftPanelLds = 75000;      // bytes of LDS a panel block may use (of 160 KiB per CU)
constexpr bool kFftPanelBases = true;    // column-pass twiddles from per-thread bases (FftStagesB)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// a * w (SIGN > 0) or a * conj(w)
template <int SIGN>
__device__ __forceinline__ double2 cmul_tw(double2 a, double2 w) {
  const double wy = SIGN > 0 ? w.y : -w.y;
  return make_double2(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
}
// a * (SIGN i)
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 a) {
  return SIGN > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// y_t = sum_s x_s exp(SIGN i 2 pi s t / R), in place, natural order in and out
template <int R, int SIGN> struct Dft;
template <int SIGN> struct Dft<2, SIGN> {
  static __device__ __forceinline__ void run(double2 (&v)[2]) {
    const double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  }
};
template <int SIGN> struct Dft<4, SIGN> {
  static __device__ __forceinline__ void run(double2 (&v)[4]) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 b0 = cadd(v[1], v[3]), b1 = mul_si<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, b0); v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1); v[3] = csub(a1, b1);
  }
};
template <int SIGN> struct Dft<8, SIGN> {
  static __device__ __forceinline__ void run(double2 (&v)[8]) {
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, SIGN>::run(e);
    Dft<4, SIGN>::run(o);
    const double h = 0.70710678118654752440;   // 1 / sqrt(2)
    // o_t exp(SIGN i 2 pi t / 8): t = 1: (1 + SIGN i) / sqrt 2, t = 2: SIGN i, t = 3: (-1 + SIGN i) / sqrt 2
    const double2 o1 = SIGN > 0 ? make_double2((o[1].x - o[1].y) * h, (o[1].y + o[1].x) * h)
                                : make_double2((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    const double2 o2 = mul_si<SIGN>(o[2]);
    const double2 o3 = SIGN > 0 ? make_double2((-o[3].x - o[3].y) * h, (-o[3].y + o[3].x) * h)
                                : make_double2((-o[3].x + o[3].y) * h, (-o[3].y - o[3].x) * h);
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

__host__ __device__ constexpr int fft_threads(int N) { return N >= 8 ? N / 8 : 1; }
__host__ __device__ constexpr int fft_lds_ix(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int fft_lds_len(int N) { return fft_lds_ix(N - 1) + 1; }

// Radix of the stage on sub-transforms of length L: 8 while the length allows, then 4 or 2.
__host__ __device__ constexpr int fft_radix(int N, int L) {
  return (N / L >= 8 && (N / fft_threads(N)) % 8 == 0) ? 8 : ((N / L >= 4 && (N / fft_threads(N)) % 4 == 0) ? 4 : 2);
}
__host__ __device__ constexpr int fft_nstages(int N) {
  int n = 0;
  for (int L = 1; L < N; L *= fft_radix(N, L)) n++;
  return n;
}

// One radix-R stage on sub-transforms of length L (see above).  v holds the thread's
// values in the t + m T layout on entry and on exit.  Every thread of the block calls
// it: a stage other than the last synchronises around its LDS exchange (lds: this
// transform's LDS, lds_stride apart for consecutive indices).  twiddles(q, k, w) puts
// w^(s k TS) into w[s], 1 <= s < R, for butterfly q (k = its index mod L).
template <int N, int R, int L, int SIGN, class Tw>
__device__ __forceinline__ void fft_reg_stage(double2 (&v)[N / fft_threads(N)], int t, const Tw& twiddles,
                                              double2* lds, int lds_stride) {
  constexpr int T = fft_threads(N), V = N / T, BQ = V / R;
  static_assert(V % R == 0 && L * R <= N, "stage does not fit the register layout");
  double2 y[BQ][R];
#pragma unroll
  for (int q = 0; q < BQ; q++) {
    const int b = t + q * T, k = b & (L - 1);
#pragma unroll
    for (int s = 0; s < R; s++) y[q][s] = v[q + s * BQ];
    if (L > 1) {
      // w^0 = 1 exactly: the k = 0 butterflies are multiplied too (exactly), no branch
      double2 w[R];
      twiddles(q, k, w);
#pragma unroll
      for (int s = 1; s < R; s++) y[q][s] = cmul_tw<SIGN>(y[q][s], w[s]);
    }
    Dft<R, SIGN>::run(y[q]);
  }
  if (L * R == N) {
#pragma unroll
    for (int q = 0; q < BQ; q++)
#pragma unroll
      for (int u = 0; u < R; u++) v[q + u * BQ] = y[q][u];
  } else {
    __syncthreads();   // every read of the previous exchange is done
#pragma unroll
    for (int q = 0; q < BQ; q++) {
      const int b = t + q * T, k = b & (L - 1), j = b / L;
#pragma unroll
      for (int u = 0; u < R; u++) lds[fft_lds_ix(j * L * R + k + u * L) * lds_stride] = y[q][u];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < V; m++) v[m] = lds[fft_lds_ix(t + m * T) * lds_stride];
  }
}
// The stages of a whole transform, from sub-transform length L on, twiddles read from the
// table tw[q] = exp(i 2 pi q / N).
template <int N, int L, int SIGN>
struct FftStages {
  static constexpr int V = N / fft_threads(N), R = fft_radix(N, L), TS = N / (L * R);
  static __device__ __forceinline__ void run(double2 (&v)[V], int t, const double2* __restrict__ tw, double2* lds,
                                             int lds_stride) {
    auto twiddles = [&](int, int k, double2 (&w)[R]) {
#pragma unroll
      for (int s = 1; s < R; s++) w[s] = tw[s * k * TS];
    };
    fft_reg_stage<N, R, L, SIGN>(v, t, twiddles, lds, lds_stride);
    if constexpr (L * R < N) FftStages<N, L * R, SIGN>::run(v, t, tw, lds, lds_stride);
  }
};
template <int N, int SIGN>
__device__ __forceinline__ void fft_reg(double2 (&v)[N / fft_threads(N)], int t, const double2* __restrict__ tw,
                                        double2* lds, int lds_stride) {
  FftStages<N, 1, SIGN>::run(v, t, tw, lds, lds_stride);
}

// The same stages with the twiddles formed from one base per butterfly and stage,
// wb[I][q] = w^(k TS) (stage I, butterfly q, k = (t + q T) mod L).  The bases depend on
// t only: a block that transforms several columns reads them once, and the stages then
// make no global memory access.  Powers by products at most three deep (w^2 = w w,
// w^3 = w^2 w, w^4 = w^2 w^2, w^5 = w^4 w, w^6 = w^4 w^2, w^7 = w^4 w^3): a few ulps
// from the table's correctly rounded values.
template <int N>
struct FftBases {
  static constexpr int NS = fft_nstages(N), WQ = (N / fft_threads(N)) / 2 > 0 ? (N / fft_threads(N)) / 2 : 1;
  double2 wb[NS][WQ];
};
template <int N, int L, int I, int SIGN>
struct FftStagesB {
  static constexpr int T = fft_threads(N), V = N / T, R = fft_radix(N, L), BQ = V / R, TS = N / (L * R);
  static __device__ __forceinline__ void load(FftBases<N>& B, int t, const GLOBAL f64x2* tw) {
#pragma unroll
    for (int q = 0; q < BQ; q++) {
      const int k = (t + q * T) & (L - 1);
      if constexpr (L > 1) {
        const f64x2 w = tw[k * TS];
        B.wb[I][q] = make_double2(w.x, w.y);
      } else {
        B.wb[I][q] = make_double2(1.0, 0.0);
      }
    }
    if constexpr (L * R < N) FftStagesB<N, L * R, I + 1, SIGN>::load(B, t, tw);
  }
  static __device__ __forceinline__ void run(double2 (&v)[V], int t, const FftBases<N>& B, double2* lds,
                                             int lds_stride) {
    auto twiddles = [&](int q, int, double2 (&w)[R]) {
      // an opaque copy of the base: the powers are formed here, in every transform,
      // rather than hoisted out of a caller's loop and held in registers
      double wx = B.wb[I][q].x, wy = B.wb[I][q].y;
      asm volatile("" : "+v"(wx), "+v"(wy));
      w[1] = make_double2(wx, wy);
      if constexpr (R >= 4) {
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[2], w[1]);
      }
      if constexpr (R == 8) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
      }
    };
    fft_reg_stage<N, R, L, SIGN>(v, t, twiddles, lds, lds_stride);
    if constexpr (L * R < N) FftStagesB<N, L * R, I + 1, SIGN>::run(v, t, B, lds, lds_stride);
  }
};

__global__ void k_twiddles(double2* tw, int n) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    double sn, cs;
    sincospi(2.0 * (double)q / (double)n, &sn, &cs);
    tw[q] = make_double2(cs, sn);
  }
}

// Along the contiguous (row) index: C = kFftBlock / T columns (n values each, T threads
// each) per block row br, column gc = br C + g of the batch (field gc / N).  The block
// rows are spread over a persistent grid (blockIdx.x, then gridDim.x apart): while a
// block transforms one column it has the next one's input in flight, the field values
// or, where `pre` generates the field, the maps of pre's first two elements when they
// are masks or mirrors.  The stages take their twiddles from per-thread bases
// (FftStagesB), so between the input and the output a column's only global accesses
// are the chains' other maps.  The field's pointwise chain `pre` (point-source
// generation, masks, mirrors: the elements before the transform) is applied as the
// column is read, `post` (the masks and mirrors after it) and `scale` as it is written.
template <int N>
__host__ __device__ constexpr int fft_contig_block() { return fft_threads(N) > kFftBlock ? fft_threads(N) : kFftBlock; }
constexpr int kFftRowWaves = 2;          // waves per SIMD a row-pass block is compiled for
// What a row pass loads ahead.  PF >= 0: the field is generated by `pre` and the next
// column's maps of pre's elements 1 .. PF (masks or mirrors, PF <= 3) are prefetched;
// PF = kPfField: the next column's field values are prefetched.  EP: the current
// column's map of post's first element (a mask or a mirror) is loaded before the
// transform rather than after it.
constexpr int kPfField = -1;
template <int N, int SIGN, int PF, bool EP>
__global__ __launch_bounds__(fft_contig_block<N>()) __attribute__((amdgpu_waves_per_eu(kFftRowWaves)))
void k_fft_contig(double2* __restrict__ E, int nbr, int ncol, double scale, const double2* __restrict__ tw,
                  PwChain pre, int has_pre, PwChain post, int has_post) {
  constexpr int T = fft_threads(N), V = N / T, C = fft_contig_block<N>() / T;
  constexpr int npre = PF > 0 ? PF : 0, NX = PF == kPfField ? 2 : (npre > 0 ? npre : 1);
  extern __shared__ double2 sbuf[];
  const int g = threadIdx.x / T, t0 = threadIdx.x - g * T;
  double2* lds = sbuf + g * fft_lds_len(N);
  FftBases<N> B;
  FftStagesB<N, 1, 0, SIGN>::load(B, t0, gin<f64x2>(tw));
  // the loads of block row br (the spare groups of the last block row read the batch's
  // last column: their values are never stored)
  auto fetch = [&](int br, int t, double (&dst)[NX][V]) {
    const int gc = min(br * C + g, ncol - 1);
    const int f = gc / N, p0 = (gc - f * N) * N + t;
    if constexpr (PF == kPfField) {
      const GLOBAL f64x2* Ef = gout<f64x2>(E) + (long)f * N * N + p0;
#pragma unroll
      for (int m = 0; m < V; m++) { const f64x2 e = Ef[m * T]; dst[0][m] = e.x; dst[1][m] = e.y; }
    } else {
#pragma unroll
      for (int q = 0; q < npre; q++) {
        const GLOBAL double* a = gin<double>(pre.op[1 + q].p) + p0;
#pragma unroll
        for (int m = 0; m < V; m++) dst[q][m] = a[m * T];
      }
    }
  };
  double nx[NX][V];
#pragma unroll
  for (int q = 0; q < NX; q++)
#pragma unroll
    for (int m = 0; m < V; m++) nx[q][m] = 0.0;
  if (blockIdx.x < nbr) fetch(blockIdx.x, t0, nx);
  for (int br = blockIdx.x; br < nbr; br += gridDim.x) {
    // an opaque copy of t per column: the addresses derived from it are formed again in
    // every iteration rather than held in registers across the loop (they would not fit)
    int t = t0;
    asm volatile("" : "+v"(t));
    const int gc = br * C + g;
    const bool live = gc < ncol;               // spare groups of the last block row still synchronise
    const int f = live ? gc / N : 0, p0 = (live ? gc - f * N : 0) * N;
    double2 v[V];
    if (has_pre) {
      const PwField F = pw_field(pre, f);
      int pp[V], rr[V];
      double er[V], ei[V];
#pragma unroll
      for (int m = 0; m < V; m++) {
        rr[m] = t + m * T;
        pp[m] = p0 + rr[m];
        er[m] = nx[0][m]; ei[m] = nx[NX - 1][m];   // the field (overwritten where it is generated)
      }
      if (live) pw_points<V, npre>(pre, F, f, pp, rr, p0 / N, er, ei, nx);
#pragma unroll
      for (int m = 0; m < V; m++) v[m] = make_double2(er[m], ei[m]);
    } else {
#pragma unroll
      for (int m = 0; m < V; m++) v[m] = make_double2(nx[0][m], nx[NX - 1][m]);
    }
    // the next block row's loads (and this column's post map), in flight during the
    // transform.  Every earlier load has been used by now; waiting for them here keeps
    // the compiler's count of loads in flight exact where the branch joins (it would
    // otherwise wait for these loads too)
    __builtin_amdgcn_s_waitcnt(0x0F70);                       // vmcnt(0)
    double pe[1][V];
    if constexpr (EP) {
      const GLOBAL double* a = gin<double>(post.op[0].p) + p0 + t;
#pragma unroll
      for (int m = 0; m < V; m++) pe[0][m] = a[m * T];
    }
    if (br + (int)gridDim.x < nbr) fetch(br + gridDim.x, t, nx);
    FftStagesB<N, 1, 0, SIGN>::run(v, t, B, lds, 1);
    if (!live) continue;
    GLOBAL f64x2* Ef = gout<f64x2>(E) + (long)f * N * N;
    if (has_post) {
      const PwField G = pw_field(post, f);
      int pp[V], rr[V];
      double er[V], ei[V];
#pragma unroll
      for (int m = 0; m < V; m++) {
        rr[m] = t + m * T;
        pp[m] = p0 + rr[m];
        er[m] = v[m].x * scale; ei[m] = v[m].y * scale;
      }
      pw_points<V, EP ? 1 : 0>(post, G, f, pp, rr, p0 / N, er, ei, pe);
#pragma unroll
      for (int m = 0; m < V; m++) Ef[pp[m]] = f64x2{er[m], ei[m]};
    } else {
#pragma unroll
      for (int m = 0; m < V; m++) Ef[p0 + t + m * T] = f64x2{v[m].x * scale, v[m].y * scale};
    }
  }
}

// Rows per panel: a power of two, at most N, at most kFftPanelMaxThreads / T, the
// panel's LDS (RP padded rows) within kFftPanelLds.
template <int N>
__host__ __device__ constexpr int fft_panel_rows() {
  int rp = 1;
  while (2 * rp <= N && 2 * rp * fft_threads(N) <= kFftPanelMaxThreads &&
         (long)2 * rp * fft_lds_len(N) * 16 <= kFftPanelLds)
    rp *= 2;
  return rp;
}

// Along the strided (column) index: a panel of RP consecutive rows of one field per
// block, RP T threads, thread (rr, t) = threadIdx.x (rr = x mod RP fastest) holding row
// r0 + rr's values t + m T.  Each global access of a wave is T-stride runs of RP values
// (16 RP contiguous bytes).  In LDS the rows are interleaved (value i of row rr at
// fft_lds_ix(i) RP + rr), so lanes of consecutive rows hit consecutive slots.  Panels
// are numbered XCD-major: the blocks one XCD receives (every 8th in dispatch order)
// take consecutive panels, so the runs that share a 128-byte line (RP = 2 at
// N = 2048: four panels) are read through one L2 at about the same time.
// MODE 0: the transform of sign SIGN.  MODE 1: forward, times the angular-spectrum
// transfer function, inverse: the Fresnel step's middle pass (the forward transform's
// results are the inverse's operands in registers).
template <int N, int MODE, int SIGN>
__global__ __launch_bounds__(fft_panel_rows<N>() * fft_threads(N)) void k_fft_panel(
    double2* __restrict__ E, const double* __restrict__ lambdas, double ap_rad, double z,
    const double2* __restrict__ tw) {
  constexpr int T = fft_threads(N), V = N / T, RP = fft_panel_rows<N>(), P = N / RP;
  extern __shared__ double2 sbuf[];
  const int bx = blockIdx.x;
  const int panel = (P % 8 == 0) ? (bx & 7) * (P / 8) + (bx >> 3) : bx;
  const int f = blockIdx.y, r0 = panel * RP;
  const int rr = threadIdx.x % RP, t = threadIdx.x / RP;
  GLOBAL f64x2* F = gout<f64x2>(E) + (long)f * N * N + r0 + rr;
  FftBases<N> B;
  if constexpr (kFftPanelBases) FftStagesB<N, 1, 0, -1>::load(B, t, gin<f64x2>(tw));
  double2 v[V];
#pragma unroll
  for (int m = 0; m < V; m++) {
    const f64x2 e = F[(long)(t + m * T) * N];
    v[m] = make_double2(e.x, e.y);
  }
  double2* lds = sbuf + rr;
  if (MODE == 0) {
    if constexpr (kFftPanelBases) FftStagesB<N, 1, 0, SIGN>::run(v, t, B, lds, RP);
    else fft_reg<N, SIGN>(v, t, tw, lds, RP);
  } else {
    if constexpr (kFftPanelBases) FftStagesB<N, 1, 0, -1>::run(v, t, B, lds, RP);
    else fft_reg<N, -1>(v, t, tw, lds, RP);
    {
#pragma clang fp contract(off)
      // csim: exp(i*k*z * sqrt(1. - 1./4.*pow(lambda/apRad, 2.)*H)), Goodman 4-20
      const double lam = lambdas[f];
      const double k = 2. * M_PI / lam;
      const double q = lam / ap_rad;
      const double c4 = 1. / 4. * (q * q);
      const int r = r0 + rr;
      const int pr = r < N / 2 ? r : r - N;      // fft_shift of regspace(-n/2, n/2-1)
#pragma unroll
      for (int m = 0; m < V; m++) {
        const int c = t + m * T;
        const int pc = c < N / 2 ? c : c - N;
        const double H = (double)pc * (double)pc + (double)pr * (double)pr;
        const double ph = k * z * sqrt(1. - c4 * H);
        double sn, cs;
        sincos_phase(ph, &sn, &cs);
        v[m] = cmul(v[m], make_double2(cs, sn));
        // one value at a time: interleaving the V phasors would not fit the 128 VGPRs a
        // 1024-thread block leaves
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // the inverse reads the forward's twiddles and LDS slots again: opaque copies of the
    // twiddle pointer and of t keep the compiler from holding the forward's loaded
    // twiddles and computed addresses in registers across H (they would not fit).  The
    // copies keep their address spaces (a generic pointer would turn the inverse's LDS
    // exchanges and twiddle reads into flat accesses, which wait on both counters).
    const GLOBAL double2* tw_inv = (const GLOBAL double2*)tw;
    int t_inv = t;
    asm volatile("" : "+s"(tw_inv), "+v"(t_inv));
    if constexpr (kFftPanelBases) FftStagesB<N, 1, 0, +1>::run(v, t_inv, B, lds, RP);
    else fft_reg<N, +1>(v, t_inv, (const double2*)tw_inv, lds, RP);
  }
  // the stores recompute their addresses (an opaque copy of the base): keeping the loads'
  // V addresses live through the transforms costs registers the transforms need
  GLOBAL f64x2* Fw = F;
  asm volatile("" : "+v"(Fw));
#pragma unroll
  for (int m = 0; m < V; m++) Fw[(long)(t + m * T) * N] = f64x2{v[m].x, v[m].y};
}

static int check_fft_n(int n) {
  if (n < 2 || n > 4096 || (n & (n - 1))) return fail(CSX_ERR_ARG, "fft: n must be a power of two in [2, 4096]");
  return CSX_OK;
}

// Twiddle tables exp(i 2 pi q / n), q < n, one per transform length, kept for the
// process lifetime (one per device).
static const double2* twiddles(int n, hipStream_t s) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, double2*> tables;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, n);
  auto it = tables.find(key);
  if (it != tables.end()) return it->second;
  double2* p = nullptr;
  if (hipMalloc(&p, sizeof(double2) * n) != hipSuccess) return nullptr;
  hipLaunchKernelGGL(k_twiddles, dim3(std::max(1, (n + 255) / 256)), dim3(256), 0, s, p, n);
  tables[key] = p;
  return p;
}

// Compute units of the current device (queried once per device).
static int device_cus(int* cus) {
  static std::mutex mu;
  static std::map<int, int> known;
  int dev = 0;
  CSX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = known.find(dev);
  if (it == known.end()) {
    int n = 0;
    CSX_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    it = known.emplace(dev, std::max(1, n)).first;
  }
  *cus = it->second;
  return CSX_OK;
}

// Dynamic LDS beyond 64 KiB has to be allowed per kernel (once per process).
template <class K>
static int allow_lds(K* kernel, size_t bytes) {
  if (bytes <= ((size_t)64 << 10)) return CSX_OK;
  CSX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes));
  return CSX_OK;
}

template <int N, int SIGN, int PF, bool EP>
static int launch_contig_pf(double2* E, int nfield, double scale, const double2* tw, hipStream_t s,
                            const PwChain* pre, const PwChain* post) {
  constexpr int B = fft_contig_block<N>(), C = B / fft_threads(N);
  const size_t lds = sizeof(double2) * C * fft_lds_len(N);
  static const int attr = allow_lds(&k_fft_contig<N, SIGN, PF, EP>, lds);
  CSX_TRY(attr);
  const int ncol = N * nfield, nbr = (ncol + C - 1) / C;
  // persistent grid: the blocks that fit on the device at once (kFftRowWaves waves per
  // SIMD, as the LDS allows), each taking every grid-th block row
  int cus = 0;
  CSX_TRY(device_cus(&cus));
  const int waves = B / 64, per_cu_w = 4 * kFftRowWaves / waves;
  const int per_cu_l = (int)((160 << 10) / lds);
  const int grid = std::min(nbr, cus * std::max(1, std::min(per_cu_w, per_cu_l)));
  PwChain none{};
  hipLaunchKernelGGL((k_fft_contig<N, SIGN, PF, EP>), dim3(grid), dim3(B), lds, s, E, nbr, ncol, scale, tw,
                     pre ? *pre : none, pre ? 1 : 0, post ? *post : none, post ? 1 : 0);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}
// The load-ahead form for a pass's chains.  A generating chain (its first element the
// point source) comes only before a forward pass, a post chain only after an inverse one
// (the Fresnel step's passes); the forms beyond the plain ones are built for N >= 512.
static bool pw_is_map(const PwChain* c, int k) {
  return c && k < c->nops && (c->op[k].kind == PW_MASK_1PI || c->op[k].kind == PW_MIRROR);
}
template <int N, int SIGN>
static int launch_contig(double2* E, int nfield, double scale, const double2* tw, hipStream_t s,
                         const PwChain* pre, const PwChain* post) {
  const bool gen = pre && pre->nops > 0 && pre->op[0].kind == PW_SOURCE;
  if constexpr (N >= 512 && SIGN < 0) {
    if (gen) {
      const int nm = pw_is_map(pre, 1) ? (pw_is_map(pre, 2) ? (pw_is_map(pre, 3) ? 3 : 2) : 1) : 0;
      if (post) return fail(CSX_ERR_ARG, "fft: a post chain after a forward transform");
      switch (nm) {
        case 3: return launch_contig_pf<N, SIGN, 3, false>(E, nfield, scale, tw, s, pre, post);
        case 2: return launch_contig_pf<N, SIGN, 2, false>(E, nfield, scale, tw, s, pre, post);
        case 1: return launch_contig_pf<N, SIGN, 1, false>(E, nfield, scale, tw, s, pre, post);
        default: return launch_contig_pf<N, SIGN, 0, false>(E, nfield, scale, tw, s, pre, post);
      }
    }
  }
  if constexpr (N >= 512 && SIGN > 0) {
    if (!gen && pw_is_map(post, 0)) return launch_contig_pf<N, SIGN, kPfField, true>(E, nfield, scale, tw, s, pre, post);
  }
  if constexpr (SIGN < 0) {
    if (gen) return launch_contig_pf<N, SIGN, 0, false>(E, nfield, scale, tw, s, pre, post);
  } else {
    if (gen) return fail(CSX_ERR_ARG, "fft: a generating chain before an inverse transform");
  }
  return launch_contig_pf<N, SIGN, kPfField, false>(E, nfield, scale, tw, s, pre, post);
}
template <int N, int MODE, int SIGN>
static int launch_panel(double2* E, int nfield, const double* lambdas, double ap_rad, double z, const double2* tw,
                        hipStream_t s) {
  constexpr int RP = fft_panel_rows<N>();
  const size_t lds = sizeof(double2) * RP * fft_lds_len(N);
  static const int attr = allow_lds(&k_fft_panel<N, MODE, SIGN>, lds);
  CSX_TRY(attr);
  hipLaunchKernelGGL((k_fft_panel<N, MODE, SIGN>), dim3(N / RP, nfield), dim3(RP * fft_threads(N)), lds, s, E,
                     lambdas, ap_rad, z, tw);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

enum FftOp { kFftForward, kFftInverse, kFftFresnel };
template <int N>
static int fft_run(double2* E, int nfield, FftOp op, const double* lambdas, double ap_rad, double z,
                   const PwChain* pre, const PwChain* post, hipStream_t s) {
  const double2* tw = twiddles(N, s);
  if (!tw) return fail(CSX_ERR_NOMEM, "fft: twiddle table allocation failed");
  const double inv = 1.0 / ((double)N * (double)N);
  if (op == kFftForward) {
    CSX_TRY((launch_contig<N, -1>(E, nfield, 1.0, tw, s, nullptr, nullptr)));
    CSX_TRY((launch_panel<N, 0, -1>(E, nfield, nullptr, 0, 0, tw, s)));
  } else if (op == kFftInverse) {
    CSX_TRY((launch_panel<N, 0, +1>(E, nfield, nullptr, 0, 0, tw, s)));
    CSX_TRY((launch_contig<N, +1>(E, nfield, inv, tw, s, nullptr, nullptr)));
  } else {
    CSX_TRY((launch_contig<N, -1>(E, nfield, 1.0, tw, s, pre, nullptr)));
    CSX_TRY((launch_panel<N, 1, 0>(E, nfield, lambdas, ap_rad, z, tw, s)));
    CSX_TRY((launch_contig<N, +1>(E, nfield, inv, tw, s, nullptr, post)));
  }
  return CSX_OK;
}
static int fft_dispatch(int n, double2* E, int nfield, FftOp op, const double* lambdas, double ap_rad, double z,
                        const PwChain* pre, const PwChain* post, hipStream_t s) {
  switch (n) {
#define CSX_FFT_N(N) \
    case N: return fft_run<N>(E, nfield, op, lambdas, ap_rad, z, pre, post, s);
    CSX_FFT_N(2) CSX_FFT_N(4) CSX_FFT_N(8) CSX_FFT_N(16) CSX_FFT_N(32) CSX_FFT_N(64) CSX_FFT_N(128)
    CSX_FFT_N(256) CSX_FFT_N(512) CSX_FFT_N(1024) CSX_FFT_N(2048) CSX_FFT_N(4096)
#undef CSX_FFT_N
    default: return fail(CSX_ERR_ARG, "fft: n must be a power of two in [2, 4096]");
  }
}

int fft2_batched(double2* E, int n, int nfield, int dir, hipStream_t s) {
  CSX_TRY(check_fft_n(n));
  if (nfield <= 0) return CSX_OK;
  return fft_dispatch(n, E, nfield, dir < 0 ? kFftForward : kFftInverse, nullptr, 0, 0, nullptr, nullptr, s);
}

int fresnel_batched(double2* E, int n, int nfield, const double* lambdas_dev, double ap_rad, double z, hipStream_t s,
                    const PwChain* pre, const PwChain* post) {
  CSX_TRY(check_fft_n(n));
  if (nfield <= 0) return CSX_OK;
  if ((pre && (pre->rows != n || pre->cols != n)) || (post && (post->rows != n || post->cols != n)))
    return fail(CSX_ERR_ARG, "fresnel: pointwise chain grid differs from the field");
  return fft_dispatch(n, E, nfield, kFftFresnel, lambdas_dev, ap_rad, z, pre, post, s);
}

// --------------------------------------------------------------- resampling --
// out(r', c') = (1-wy)[(1-wx) in(ylo,xlo) + wx in(ylo,xhi)] + wy[(1-wx) in(yhi,xlo) + wx in(yhi,xhi)]
// (interp along x first, then y, as src/lib/csim_geom.cpp:244-266).
__global__ void k_resample(const double2* __restrict__ in, int rows, int cols, int nfield,
                           double2* __restrict__ out, int orows, int ocols,
                           const InterpTable* __restrict__ ty, const InterpTable* __restrict__ tx) {
#pragma clang fp contract(off)
  long onpix = (long)orows * ocols, total = onpix * nfield;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / onpix, p = i - f * onpix;
    int c = (int)(p / orows), r = (int)(p - (long)c * orows);
    InterpTable X = tx[c], Y = ty[r];
    double2 v = make_double2(0, 0);
    if (X.lo >= 0 && Y.lo >= 0) {
      const double2* F = in + f * (long)rows * cols;
      double2 a = F[(long)X.lo * rows + Y.lo], b = F[(long)X.hi * rows + Y.lo];
      double2 c2 = F[(long)X.lo * rows + Y.hi], d = F[(long)X.hi * rows + Y.hi];
      double s1r = (1.0 - X.w) * a.x + X.w * b.x, s1i = (1.0 - X.w) * a.y + X.w * b.y;
      double s2r = (1.0 - X.w) * c2.x + X.w * d.x, s2i = (1.0 - X.w) * c2.y + X.w * d.y;
      v = make_double2((1.0 - Y.w) * s1r + Y.w * s2r, (1.0 - Y.w) * s1i + Y.w * s2i);
    }
    out[i] = v;
  }
}
int resample_batched(const double2* in, int rows, int cols, int nfield, double2* out, int orows,
                     int ocols, const InterpTable* ty_dev, const InterpTable* tx_dev, hipStream_t s) {
  long total = (long)orows * ocols * nfield;
  hipLaunchKernelGGL(k_resample, dim3(grid_for(total)), dim3(256), 0, s, in, rows, cols, nfield, out, orows,
                     ocols, ty_dev, tx_dev);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ------------------------------------------------------------------ hex mask --
__global__ void k_hex(const int32_t* __restrict__ nsub, const int32_t* __restrict__ hexnum, int S,
                      int nsubpix, const double* __restrict__ sags, int nhex, int ncand,
                      const double* __restrict__ lambdas, int nlam, double phase_sign, int use_only,
                      double2* __restrict__ out) {
#pragma clang fp contract(off)
  long plane = (long)S * S, total = plane * ncand * nlam;
  double norm = (double)nsubpix * nsubpix;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long cl = i / plane, p = i - cl * plane;
    int cand = (int)(cl / nlam), li = (int)(cl - (long)cand * nlam);
    double lam = lambdas[li];
    const double* sg = sags + (long)cand * nhex;
    const int only = use_only == kUseOnlyPerCandidate ? cand : use_only;
    double ar = 0, ai = 0;
#pragma unroll
    for (int s = 0; s < 3; s++) {
      int h = hexnum[p + s * plane];
      double ns = (double)nsub[p + s * plane];
      if (only > -1) {
        if (h == only) ;
        else if (only == nhex && h == -1) ;
        else ns = 0;
      }
      double sag = h > -1 ? sg[h] : 0.0;
      // csim: exp(phaseSign * 2 * M_PI * i * 2.0 * sag / lambda)
      double ph = phase_sign * 2 * M_PI * 2.0 * sag / lam;
      double sn, cs;
      sincos_phase(ph, &sn, &cs);
      ar += ns * cs;
      ai += ns * sn;
    }
    out[i] = make_double2(ar / norm, ai / norm);
  }
}
int hex_masks(const int32_t* nsub, const int32_t* hexnum, int S, int nsubpix, const double* sags,
              int nhex, int ncand, const double* lambdas_dev, int nlam, double phase_sign,
              int use_only, double2* out, hipStream_t s) {
  long total = (long)S * S * ncand * nlam;
  hipLaunchKernelGGL(k_hex, dim3(grid_for(total)), dim3(256), 0, s, nsub, hexnum, S, nsubpix, sags, nhex,
                     ncand, lambdas_dev, nlam, phase_sign, use_only, out);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// --------------------------------------------------------------- reductions --
__global__ void k_intensity(const double2* __restrict__ E, long npix, int ngroups, int per, double* __restrict__ I) {
  long total = npix * ngroups;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long g = i / npix, p = i - g * npix;
    double acc = 0;
    for (int f = 0; f < per; f++) {
      double2 e = E[(g * per + f) * npix + p];
      acc += e.x * e.x + e.y * e.y;
    }
    I[i] = acc;
  }
}
int intensity_groups(const double2* E, long npix, int ngroups, int per, double* I, hipStream_t s) {
  long total = npix * ngroups;
  hipLaunchKernelGGL(k_intensity, dim3(grid_for(total)), dim3(256), 0, s, E, npix, ngroups, per, I);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

__device__ double block_reduce(double v, int op_max) {
  __shared__ double red[32];
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = op_max ? fmax(v, w) : v + w;
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    double r = red[0];
    for (int i = 1; i < nw; i++) r = op_max ? fmax(r, red[i]) : r + red[i];
    red[0] = r;
  }
  __syncthreads();
  return red[0];
}

__global__ void k_max(const double* __restrict__ I, long npix, double* __restrict__ out) {
  const double* x = I + blockIdx.x * npix;
  double v = -INFINITY;
  for (long p = threadIdx.x; p < npix; p += blockDim.x) v = fmax(v, x[p]);
  v = block_reduce(v, 1);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}
int max_groups(const double* I, long npix, int ngroups, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_max, dim3(ngroups), dim3(256), 0, s, I, npix, out);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// Fixed-order (deterministic) region sum: each thread sums a strided subset.
__global__ void k_region_mean(const double* __restrict__ I, long npix, const int32_t* __restrict__ idx,
                              int nidx, const double* __restrict__ peak_dev, double peak,
                              double* __restrict__ out) {
  const double* x = I + blockIdx.x * npix;
  double v = 0;
  for (int q = threadIdx.x; q < nidx; q += blockDim.x) v += x[idx[q]];
  v = block_reduce(v, 0);
  if (threadIdx.x == 0) {
    double pk = peak_dev ? peak_dev[0] : peak;
    out[blockIdx.x] = (v / nidx) / pk;
  }
}
// Contrast curve (contrastCurve::make_contrast_curve, src/tools/contrastCurve.cpp:124-133):
// bin b is the mean of I over the pixels whose radius r = sqrt(x^2 + y^2) / loD lies
// strictly between s[b] and s[b + 2], divided by the peak; NaN for an empty annulus.
__global__ void k_annulus_means(const double* __restrict__ I, int nx, int ny, double x0, double dx, double y0,
                                double dy, double loD, const double* __restrict__ smp,
                                const double* __restrict__ peak_dev, double* __restrict__ out) {
  const double lo = smp[blockIdx.x], hi = smp[blockIdx.x + 2];
  double v = 0, n = 0;
  for (int p = threadIdx.x; p < nx * ny; p += blockDim.x) {
    const int c = p / ny, r = p - c * ny;
    const double XX = x0 + c * dx, YY = y0 + r * dy;
    const double rr = sqrt(XX * XX + YY * YY) / loD;
    if (rr > lo && rr < hi) { v += I[p]; n += 1; }
  }
  v = block_reduce(v, 0);
  __syncthreads();
  n = block_reduce(n, 0);
  if (threadIdx.x == 0) out[blockIdx.x] = n > 0 ? (v / n) / peak_dev[0] : NAN;
}
int annulus_means(const double* I, int nx, int ny, double x0, double dx, double y0, double dy, double loD,
                  const double* samples_dev, int nbins, const double* peak_dev, double* out, hipStream_t s) {
  if (nbins <= 0) return CSX_OK;
  hipLaunchKernelGGL(k_annulus_means, dim3(nbins), dim3(256), 0, s, I, nx, ny, x0, dx, y0, dy, loD, samples_dev,
                     peak_dev, out);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

int region_mean_groups(const double* I, long npix, int ngroups, const int32_t* idx, int nidx,
                       const double* peak_dev, double peak, double* out, hipStream_t s) {
  if (nidx <= 0) return fail(CSX_ERR_ARG, "region_mean: empty region");
  hipLaunchKernelGGL(k_region_mean, dim3(ngroups), dim3(256), 0, s, I, npix, idx, nidx, peak_dev, peak, out);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ------------------------------------------------------------ bin profiles --
// out[f][b] = sum over pixels p with bin[p] == b of |E_f[p]|^2, fixed summation
// order (per-thread partial sums, then a sequential sum over threads).
constexpr int kBinThreads = 128;
__global__ __launch_bounds__(kBinThreads) void k_bin_sums(const double2* __restrict__ E, long npix,
                                                          const int32_t* __restrict__ bin, int nbins,
                                                          double* __restrict__ out) {
  extern __shared__ double part[];            // [kBinThreads][nbins]
  const double2* F = E + blockIdx.x * npix;
  double* mine = part + threadIdx.x * nbins;
  for (int b = 0; b < nbins; b++) mine[b] = 0.0;
  for (long p = threadIdx.x; p < npix; p += blockDim.x) {
    int b = bin[p];
    if (b >= 0 && b < nbins) {
      double2 e = F[p];
      mine[b] += e.x * e.x + e.y * e.y;
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    double acc = 0;
    for (int t = 0; t < (int)blockDim.x; t++) acc += part[t * nbins + b];
    out[(long)blockIdx.x * nbins + b] = acc;
  }
}
int bin_sums(const double2* E, long npix, int nfield, const int32_t* bin, int nbins, double* out, hipStream_t s) {
  if (nbins <= 0 || nbins > 64) return fail(CSX_ERR_ARG, "bin_sums: nbins must be in [1, 64]");
  size_t lds = sizeof(double) * kBinThreads * nbins;
  hipLaunchKernelGGL(k_bin_sums, dim3(nfield), dim3(kBinThreads), lds, s, E, npix, bin, nbins, out);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ------------------------------------------------------------------- embed --
__global__ void k_embed(const double2* __restrict__ src, int srows, int scols, double2* __restrict__ dst,
                        int drows, int dcols, int nfield, int crop) {
  // crop == 0: dst larger, src placed at (drows/2 - srows/2, dcols/2 - scols/2), zeros around.
  // crop == 1: dst smaller, dst = src window at (srows/2 - drows/2, scols/2 - dcols/2).
  long dn = (long)drows * dcols, total = dn * nfield;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / dn, p = i - f * dn;
    int c = (int)(p / drows), r = (int)(p - (long)c * drows);
    double2 v = make_double2(0, 0);
    if (!crop) {
      int r0 = drows / 2 - srows / 2, c0 = dcols / 2 - scols / 2;
      int sr = r - r0, sc = c - c0;
      if (sr >= 0 && sr < srows && sc >= 0 && sc < scols) v = src[f * (long)srows * scols + (long)sc * srows + sr];
    } else {
      int r0 = srows / 2 - drows / 2, c0 = scols / 2 - dcols / 2;
      v = src[f * (long)srows * scols + (long)(c + c0) * srows + (r + r0)];
    }
    dst[i] = v;
  }
}
int embed_center(const double2* src, int srows, int scols, double2* dst, int drows, int dcols,
                 int nfield, int crop, hipStream_t s) {
  long total = (long)drows * dcols * nfield;
  hipLaunchKernelGGL(k_embed, dim3(grid_for(total)), dim3(256), 0, s, src, srows, scols, dst, drows, dcols,
                     nfield, crop);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ------------------------------------------------ convolution-form FPM --
// fpmPupToLyot (src/elements/fpmPupToLyot.cpp:24-167).  Kernel input of the mask
// transform: out[k] = factor * (src[k] - minus1), columns mirrored for fliplr.
__global__ void k_fpm_prep(const double2* __restrict__ src, long src_stride, int rows, int cols, int n,
                           int fliplr, int minus1_from, double factor, double2* __restrict__ out) {
  long np = (long)rows * cols, total = np * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long k = i / np, p = i - k * np;
    int c = (int)(p / rows), r = (int)(p - (long)c * rows);
    int sc = fliplr ? cols - 1 - c : c;
    double2 v = src[k * src_stride + (long)sc * rows + r];
    if (k >= minus1_from) v.x -= 1.0;
    out[i] = make_double2(factor * v.x, factor * v.y);
  }
}
int fpm_prep(const double2* src, long src_stride, int rows, int cols, int n, int fliplr, int minus1_from,
             double factor, double2* out, hipStream_t s) {
  long total = (long)rows * cols * n;
  if (total <= 0) return CSX_OK;
  hipLaunchKernelGGL(k_fpm_prep, dim3(grid_for(total)), dim3(256), 0, s, src, src_stride, rows, cols, n, fliplr,
                     minus1_from, factor, out);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// P[f](r, c) *= H[h(f)](r, c) (-1)^(r + c): the transform of fft_shift(maskHat)
// for an even size is (-1)^(r+c) times the transform of maskHat.
// P[f] = H[cand(f), f % nl] * Pin[f % nin] * (-1)^(r+c) (fft_shift of the kernel folded
// in); in place when Pin == P and nin == the field count.
__global__ void k_spectrum_mul(double2* __restrict__ P, const double2* __restrict__ Pin, int nin, int n2,
                               const double2* __restrict__ H, int per_cand, int nl, int nmask_cand) {
  const int f = blockIdx.y;
  int cand = f / per_cand, l = f % nl;
  if (cand >= nmask_cand) cand = nmask_cand - 1;
  const long np = (long)n2 * n2;
  double2* p = P + (long)f * np;
  const double2* q = Pin + (long)(f % nin) * np;
  const double2* h = H + ((long)cand * nl + l) * np;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < np; i += (long)gridDim.x * blockDim.x) {
    int c = (int)(i / n2), r = (int)(i - (long)c * n2);
    double2 a = h[i], b = q[i];
    double2 v = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    if ((r + c) & 1) { v.x = -v.x; v.y = -v.y; }
    p[i] = v;
  }
}
int spectrum_mul(double2* P, const double2* Pin, int nin, int n2, int nfield, const double2* H, int per_cand, int nl,
                 int nmask_cand, hipStream_t s) {
  if (nfield <= 0) return CSX_OK;
  if (nin <= 0 || (Pin == P && nin != nfield)) return fail(CSX_ERR_ARG, "spectrum_mul: in-place needs every field");
  long np = (long)n2 * n2;
  int bx = (int)std::min<long>(2048, (np + 255) / 256);
  hipLaunchKernelGGL(k_spectrum_mul, dim3(bx, nfield), dim3(256), 0, s, P, Pin, nin, n2, H, per_cand, nl,
                     nmask_cand);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// E[f] = scale[l] * P[f](centre) + (babinet ? Ein[f % nin] : 0), centre rows/cols n/2 .. 3n/2 - 1;
// in place when Ein == E and nin == the field count.
__global__ void k_crop_add(double2* __restrict__ E, const double2* __restrict__ Ein, int nin, int n,
                           const double2* __restrict__ P, int n2, int nfield, int nl,
                           const double2* __restrict__ scale, int babinet_from) {
  long np = (long)n * n, total = np * nfield;
  int o = n / 2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / np, p = i - f * np;
    int c = (int)(p / n), r = (int)(p - (long)c * n);
    double2 a = P[f * (long)n2 * n2 + (long)(c + o) * n2 + (r + o)];
    double2 sc = scale[f % nl];
    double2 v = make_double2(a.x * sc.x - a.y * sc.y, a.x * sc.y + a.y * sc.x);
    if (f >= babinet_from) { double2 e = Ein[(f % nin) * np + p]; v.x += e.x; v.y += e.y; }
    E[i] = v;
  }
}
int crop_add(double2* E, const double2* Ein, int nin, int n, const double2* P, int n2, int nfield, int nl,
             const double2* scale_dev, int babinet_from, hipStream_t s) {
  long total = (long)n * n * nfield;
  if (total <= 0) return CSX_OK;
  if (nin <= 0 || (Ein == E && nin != nfield)) return fail(CSX_ERR_ARG, "crop_add: in-place needs every field");
  hipLaunchKernelGGL(k_crop_add, dim3(grid_for(total)), dim3(256), 0, s, E, Ein, nin, n, P, n2, nfield, nl, scale_dev,
                     babinet_from);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}


// Response-matrix gather (makeHexCFpmResponse.cpp:144-184): field f = (c * nsrc + s) *
// nl + l, region pixel j -> M[slice = s * nl + l][column c][j] (column-major npix x ncol).
__global__ void k_gather_response(const double2* __restrict__ E, long npix_field, const int32_t* __restrict__ idx,
                                  int nidx, int ncol, int nslice, double2* __restrict__ M) {
  long total = (long)nidx * ncol * nslice;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long j = i % nidx, rest = i / nidx;
    int c = (int)(rest % ncol), sl = (int)(rest / ncol);
    long f = (long)c * nslice + sl;
    M[i] = E[f * npix_field + idx[j]];
  }
}
int gather_response(const double2* E, long npix_field, const int32_t* idx, int nidx, int ncol, int nslice,
                    double2* M, hipStream_t s) {
  long total = (long)nidx * ncol * nslice;
  if (total <= 0) return CSX_OK;
  hipLaunchKernelGGL(k_gather_response, dim3(grid_for(total)), dim3(256), 0, s, E, npix_field, idx, nidx, ncol,
                     nslice, M);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

}  // namespace csx

namespace csx {

// ------------------------------------------------------------ Babinet window --
// (M - 1) E on the hex window: the part of a candidate's FPM-plane field that
// differs from the shared, unmasked one (the mask is 1 outside the window).
__global__ void k_babinet_window(const void* __restrict__ E, int kind, int terms, long e_npix, int rows,
                                 const double2* __restrict__ masks, int S, int row0, int col0, int nl, int ns,
                                 int ncm, void* __restrict__ W) {
  const int f = blockIdx.y;
  const int l = f % nl, s = (f / nl) % ns, cand = f / (nl * ns);
  const long fin = (long)s * nl + l;
  const long wn = (long)S * S;
  const GLOBAL f64x2* M = gin<f64x2>(masks) + ((long)min(cand, ncm - 1) * nl + l) * wn;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < (int)wn; p += gridDim.x * blockDim.x) {
    const int c = p / S, r = p - c * S;
    const long src = (long)(col0 + c) * rows + (row0 + r);
    double er, ei;
    if (kind == 0) {
      const f64x2 v = gin<f64x2>(E)[fin * e_npix + src];
      er = v.x; ei = v.y;
    } else if (kind == 1) {
      const f32x2 v = gin<f32x2>(E)[fin * e_npix + src];
      er = v.x; ei = v.y;
    } else {
      const GLOBAL unsigned short* e = gin<unsigned short>(E) + fin * 2 * terms * e_npix;
      er = 0.0; ei = 0.0;
      for (int t = 0; t < terms; t++) {
        er += (double)bf16_val(e[(2 * t) * e_npix + src]);
        ei += (double)bf16_val(e[(2 * t + 1) * e_npix + src]);
      }
    }
    const f64x2 m = M[p];
    const double mr = m.x - 1.0, mi = m.y;
    const double wr = mr * er - mi * ei, wi = mr * ei + mi * er;
    if (kind == 0) {
      gout<f64x2>(W)[f * wn + p] = f64x2{wr, wi};
    } else if (kind == 1) {
      gout<f32x2>(W)[f * wn + p] = f32x2{(float)wr, (float)wi};
    } else {
      // three-term planes: windows of 32 pixels of a column (S is a multiple of 32)
      int e = 0;
      if (terms == 3) e = window_exponent(window_max<kPlaneWindow>((float)fmax(fabs(wr), fabs(wi))));
      store_planes_w(gout<unsigned short>(W) + (long)f * 2 * terms * wn, p, wn, terms, wr, wi, e);
    }
  }
}

int babinet_window(const void* E, int kind, int terms, long e_npix, int rows, const double2* masks, int S,
                   int row0, int col0, int nfield, int nl, int ns, int ncm, void* W, hipStream_t s) {
  if (nfield <= 0) return CSX_OK;
  if (S <= 0 || row0 < 0 || col0 < 0 || nl <= 0 || ns <= 0 || ncm <= 0 || (kind == 2 && terms <= 0))
    return fail(CSX_ERR_ARG, "babinet_window: bad shape");
  if (kind == 2 && terms == 3 && S % kPlaneWindow)
    return fail(CSX_ERR_UNSUPPORTED, "babinet_window: three-term planes need a window size divisible by 32");
  const int blocks = std::min(64, (S * S + 255) / 256);
  hipLaunchKernelGGL(k_babinet_window, dim3(blocks, nfield), dim3(256), 0, s, E, kind, terms, e_npix, rows,
                     reinterpret_cast<const double2*>(masks), S, row0, col0, nl, ns, ncm, W);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

}  // namespace csx
