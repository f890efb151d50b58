// Grouped complex GEMM kernels on CDNA4 (gfx950) MFMA and DFT-kernel generation.
//
// Every matrix Fourier transform of the path (csim zoomFft, src/lib/csim_fft.cpp:
// 670-826, which csim evaluates with a Bluestein chirp-z transform) is computed in
// its closed form as two complex GEMMs.  One launch covers every field of a batch
// (wavelength x source x candidate, uneven sizes allowed): a grouped GEMM whose
// workgroups find their problem by binary search over per-problem tile offsets.
#include "gemm.hpp"

#include <hip/hip_bf16.h>
#include <algorithm>
#include <vector>

namespace csx {

// ------------------------------------------------------------ problem plan --
int plan_tiles(GemmProb* probs, int n, int tile_m, int tile_n) {
  int total = 0;
  for (int i = 0; i < n; i++) {
    probs[i].tiles_m = (probs[i].M + tile_m - 1) / tile_m;
    probs[i].tiles_n = (probs[i].N + tile_n - 1) / tile_n;
    probs[i].tile_start = total;
    total += probs[i].tiles_m * probs[i].tiles_n;
  }
  return total;
}

// MI355X hands consecutive workgroups of a launch to its 8 XCDs round-robin.
// Remap so that each XCD (own L2) works a contiguous range of logical tiles:
// tiles of one field share operands and stay in one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  int xcd = b & 7, q = nb >> 3, r = nb & 7;
  int start = xcd * q + (xcd < r ? xcd : r);
  return start + (b >> 3);
}

__device__ __forceinline__ int find_prob(const GemmProb* __restrict__ probs, int n, int tile) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile_start <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(r >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// (re, im) split into `terms` BF16 parts each: x = hi + mid + lo, every part the
// round-to-nearest-even of what the previous ones leave (from the FP32 value).
__device__ __forceinline__ void store_planes(unsigned short* base, long off, long stride, int terms,
                                             double vr, double vi) {
  float xr = (float)vr, xi = (float)vi;
  for (int t = 0; t < terms; t++) {
    unsigned short hr = bf16_rne(xr), hi = bf16_rne(xi);
    base[(2 * t) * stride + off] = hr;
    base[(2 * t + 1) * stride + off] = hi;
    xr -= bf16_val(hr);
    xi -= bf16_val(hi);
  }
}

// Epilogue: alpha, mask, store (shared by every kernel).
__device__ __forceinline__ void epilogue_store(const GemmProb& P, int m, int n, double re, double im) {
  if (m >= P.M || n >= P.N) return;
  double vr = P.alpha_re * re - P.alpha_im * im;
  double vi = P.alpha_re * im + P.alpha_im * re;
  if (P.mask_kind != MASK_NONE) {
    int mr = m - P.mm0, nc = n - P.mn0;
    if (mr >= 0 && mr < P.mM && nc >= 0 && nc < P.mN) {
      long off = (long)mr * P.ldm + nc;
      if (P.mask_kind == MASK_COMPLEX) {
        double2 w = static_cast<const double2*>(P.mask)[off];
        double tr = vr * w.x - vi * w.y;
        vi = vr * w.y + vi * w.x;
        vr = tr;
      } else {
        double w = static_cast<const double*>(P.mask)[off];
        if (P.mask_kind == MASK_REAL_1PI) {   // csim mask: (1+i) w
          double tr = (vr - vi) * w;
          vi = (vr + vi) * w;
          vr = tr;
        } else {
          vr *= w; vi *= w;
        }
      }
    }
  }
  long off = P.transC ? (long)n * P.ldc + m : (long)m * P.ldc + n;
  if (P.out_kind == OUT_C128) static_cast<double2*>(P.C)[off] = make_double2(vr, vi);
  else if (P.out_kind == OUT_C64) static_cast<float2*>(P.C)[off] = make_float2((float)vr, (float)vi);
  else store_planes(static_cast<unsigned short*>(P.C), off, P.sC, P.terms, vr, vi);
}

// ------------------------------------------------------ FP64 / FP32 MFMA GEMM --
// Accumulator layout (measured on gfx950): value v of lane l of the FP64
// 16x16x4 MFMA is D[4v + l/16][l%16]; the FP32 16x16x4 one is D[4(l/16) + v][l%16].
template <typename T> struct Mma;
template <> struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double2 c2;
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float2 c2;
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// 64x64 output tile, 256 threads = 4 waves of 32x32 (2x2 MFMA 16x16 sub-tiles).
// K staged through LDS 16 complex values at a time, next slab prefetched in registers.
// MFMA 16x16x4 operand layout: lane l holds A[l%16][l/16] and B[l/16][l%16].
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_nt(const GemmProb* __restrict__ probs, int nprob) {
  typedef typename Mma<T>::c2 C2;
  typedef typename Mma<T>::v4 V4;
  constexpr int BK = 16, LDK = BK + 1;
  __shared__ C2 As[kTileM][LDK];
  __shared__ C2 Bs[kTileN][LDK];

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const GemmProb& P = probs[find_prob(probs, nprob, tile)];
  const int local = tile - P.tile_start;
  const int tm = local % P.tiles_m, tn = local / P.tiles_m;
  const int m0 = tm * kTileM, n0 = tn * kTileN;
  const C2* __restrict__ A = static_cast<const C2*>(P.A);
  const C2* __restrict__ B = static_cast<const C2*>(P.B);

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = t >> 2, lk = (t & 3) * 4;   // loader: 64 rows x 4 chunks of 4

  V4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) { accR[i][j] = V4{0, 0, 0, 0}; accI[i][j] = V4{0, 0, 0, 0}; }

  C2 ra[4], rb[4];
  auto gload = [&](int k0) {
    int am = m0 + lrow, bn = n0 + lrow;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k = k0 + lk + q;
      ra[q] = (am < P.M && k < P.K) ? A[(long)am * P.lda + k] : C2{0, 0};
      rb[q] = (bn < P.N && k < P.K) ? B[(long)bn * P.ldb + k] : C2{0, 0};
    }
  };
  gload(0);
  for (int k0 = 0; k0 < P.K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) { As[lrow][lk + q] = ra[q]; Bs[lrow][lk + q] = rb[q]; }
    __syncthreads();
    if (k0 + BK < P.K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      C2 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) a[i] = As[wm * 32 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; j++) b[j] = Bs[wn * 32 + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          accR[i][j] = Mma<T>::mma(a[i].x, b[j].x, accR[i][j]);
          accI[i][j] = Mma<T>::mma(a[i].x, b[j].y, accI[i][j]);
          accR[i][j] = Mma<T>::mma(-a[i].y, b[j].y, accR[i][j]);
          accI[i][j] = Mma<T>::mma(a[i].y, b[j].x, accI[i][j]);
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int m = m0 + wm * 32 + i * 16 + Mma<T>::row(lane, v);
        int n = n0 + wn * 32 + j * 16 + (lane & 15);
        epilogue_store(P, m, n, (double)accR[i][j][v], (double)accI[i][j][v]);
      }
}

int gemm_f64(const GemmProb* probs_dev, int n, int total_tiles, hipStream_t s) {
  if (total_tiles <= 0) return CSX_OK;
  hipLaunchKernelGGL(k_gemm_nt<double>, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}
int gemm_f32(const GemmProb* probs_dev, int n, int total_tiles, hipStream_t s) {
  if (total_tiles <= 0) return CSX_OK;
  hipLaunchKernelGGL(k_gemm_nt<float>, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ------------------------------------------------------- split-BF16 MFMA GEMM --
// FP32 complex operands are split on the fly into TERMS bf16 planes
// (x = hi + mid + lo, each round-to-nearest-even of the remainder) and
// contracted with v_mfma_f32_16x16x32_bf16.  Products kept: every (a, b) term
// pair with a + b < TERMS (3 for two terms, 6 for three).  FP32 accumulators
// are flushed into FP64 after every K slab of 32 when TERMS == 3.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne_bits(float x) { return bf16_rne(x); }
__device__ __forceinline__ float bf16_bits_to_float(unsigned short h) { return bf16_val(h); }

template <int TERMS>
__global__ __launch_bounds__(256) void k_gemm_bf16split(const GemmProb* __restrict__ probs, int nprob) {
  constexpr int BK = 32;          // one MFMA K step
  constexpr int LDK = BK + 8;     // bf16 elements per LDS row (16 B pad)
  // planes: [term][re/im][row][k]
  __shared__ unsigned short As[TERMS][2][kTileM][LDK];
  __shared__ unsigned short Bs[TERMS][2][kTileN][LDK];

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const GemmProb& P = probs[find_prob(probs, nprob, tile)];
  const int local = tile - P.tile_start;
  const int tm = local % P.tiles_m, tn = local / P.tiles_m;
  const int m0 = tm * kTileM, n0 = tn * kTileN;
  const float2* __restrict__ A = static_cast<const float2*>(P.A);
  const float2* __restrict__ B = static_cast<const float2*>(P.B);

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = t >> 2, lk = (t & 3) * 8;   // loader: 64 rows x 4 chunks of 8

  f32x4 accR[2][2], accI[2][2];
  double dR[2][2][4], dI[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      accR[i][j] = f32x4{0, 0, 0, 0}; accI[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) { dR[i][j][v] = 0; dI[i][j][v] = 0; }
    }

  float2 ra[8], rb[8];
  auto gload = [&](int k0) {
    int am = m0 + lrow, bn = n0 + lrow;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int k = k0 + lk + q;
      ra[q] = (am < P.M && k < P.K) ? A[(long)am * P.lda + k] : make_float2(0.f, 0.f);
      rb[q] = (bn < P.N && k < P.K) ? B[(long)bn * P.ldb + k] : make_float2(0.f, 0.f);
    }
  };
  auto split_store = [&](unsigned short (*S)[2][kTileM][LDK], const float2* r) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      float xr = r[q].x, xi = r[q].y;
#pragma unroll
      for (int tt = 0; tt < TERMS; tt++) {
        unsigned short hr = bf16_rne_bits(xr), hi = bf16_rne_bits(xi);
        S[tt][0][lrow][lk + q] = hr;
        S[tt][1][lrow][lk + q] = hi;
        xr -= bf16_bits_to_float(hr);
        xi -= bf16_bits_to_float(hi);
      }
    }
  };

  gload(0);
  for (int k0 = 0; k0 < P.K; k0 += BK) {
    __syncthreads();
    split_store(As, ra);
    split_store(Bs, rb);
    __syncthreads();
    if (k0 + BK < P.K) gload(k0 + BK);
    // fragments: lane l holds row l%16, k = 8*(l/16) .. +7
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    u32x4 a[TERMS][2][2], b[TERMS][2][2];   // [term][re/im][subtile]
#pragma unroll
    for (int tt = 0; tt < TERMS; tt++)
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int s = 0; s < 2; s++) {
          a[tt][c][s] = *reinterpret_cast<const u32x4*>(&As[tt][c][wm * 32 + s * 16 + fr][fk]);
          b[tt][c][s] = *reinterpret_cast<const u32x4*>(&Bs[tt][c][wn * 32 + s * 16 + fr][fk]);
        }
#pragma unroll
    for (int ta = 0; ta < TERMS; ta++)
#pragma unroll
      for (int tb = 0; tb < TERMS - ta; tb++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            bf16x8 ar = __builtin_bit_cast(bf16x8, a[ta][0][i]);
            bf16x8 ai = __builtin_bit_cast(bf16x8, a[ta][1][i]);
            bf16x8 ain = __builtin_bit_cast(bf16x8, a[ta][1][i] ^ u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u});
            bf16x8 br = __builtin_bit_cast(bf16x8, b[tb][0][j]);
            bf16x8 bi = __builtin_bit_cast(bf16x8, b[tb][1][j]);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, br, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, bi, accI[i][j], 0, 0, 0);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ain, bi, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai, br, accI[i][j], 0, 0, 0);
          }
    if (TERMS >= 3) {
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
#pragma unroll
          for (int v = 0; v < 4; v++) { dR[i][j][v] += accR[i][j][v]; dI[i][j][v] += accI[i][j][v]; }
          accR[i][j] = f32x4{0, 0, 0, 0}; accI[i][j] = f32x4{0, 0, 0, 0};
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
        int n = n0 + wn * 32 + j * 16 + (lane & 15);
        double re = dR[i][j][v] + (double)accR[i][j][v];
        double im = dI[i][j][v] + (double)accI[i][j][v];
        epilogue_store(P, m, n, re, im);
      }
}

int gemm_bf16split(const GemmProb* probs_dev, int n, int total_tiles, int terms, hipStream_t s) {
  if (total_tiles <= 0) return CSX_OK;
  if (terms == 2)
    hipLaunchKernelGGL(k_gemm_bf16split<2>, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
  else if (terms == 3)
    hipLaunchKernelGGL(k_gemm_bf16split<3>, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
  else
    return fail(CSX_ERR_ARG, "gemm_bf16split: terms must be 2 or 3");
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ------------------------------------------------- BF16-plane MFMA GEMM --
// Both operands arrive as BF16 planes (split once: DFT kernels at plan time,
// fields by the producing epilogue), so the main loop is 16-byte copies into
// LDS and v_mfma_f32_16x16x32_bf16; products kept: (a, b) terms with a + b < TERMS.
// FP32 accumulators are flushed into FP64 after every K slab of 32 when TERMS == 3.
template <int TERMS>
__global__ __launch_bounds__(256) void k_gemm_planes(const GemmProb* __restrict__ probs, int nprob) {
  constexpr int BK = 32, LDK = BK + 8, NP = 2 * TERMS;
  __shared__ __align__(16) unsigned short As[NP][kTileM][LDK];
  __shared__ __align__(16) unsigned short Bs[NP][kTileN][LDK];

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const GemmProb& P = probs[find_prob(probs, nprob, tile)];
  const int local = tile - P.tile_start;
  const int tm = local % P.tiles_m, tn = local / P.tiles_m;
  const int m0 = tm * kTileM, n0 = tn * kTileN;
  const unsigned short* __restrict__ A = static_cast<const unsigned short*>(P.A);
  const unsigned short* __restrict__ B = static_cast<const unsigned short*>(P.B);

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = t >> 2, lk = (t & 3) * 8;     // loader: row, 8-element chunk

  f32x4 accR[2][2], accI[2][2];
  double dR[2][2][4], dI[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      accR[i][j] = f32x4{0, 0, 0, 0}; accI[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) { dR[i][j][v] = 0; dI[i][j][v] = 0; }
    }

  u32x4 ra[NP], rb[NP];
  auto load_chunk = [&](const unsigned short* X, long ld, long sp, int rows, int row, int k0, int q) -> u32x4 {
    u32x4 v = u32x4{0, 0, 0, 0};
    int k = k0 + lk;
    if (row < rows) {
      const unsigned short* p = X + q * sp + (long)row * ld + k;
      if (k + 8 <= P.K) {
        v = *reinterpret_cast<const u32x4*>(p);
      } else {
        unsigned short h[8];
#pragma unroll
        for (int e = 0; e < 8; e++) h[e] = (k + e < P.K) ? p[e] : (unsigned short)0;
        v = u32x4{h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16), h[4] | ((unsigned)h[5] << 16),
                  h[6] | ((unsigned)h[7] << 16)};
      }
    }
    return v;
  };
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NP; q++) {
      ra[q] = load_chunk(A, P.lda, P.sA, P.M, m0 + lrow, k0, q);
      rb[q] = load_chunk(B, P.ldb, P.sB, P.N, n0 + lrow, k0, q);
    }
  };

  gload(0);
  for (int k0 = 0; k0 < P.K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NP; q++) {
      *reinterpret_cast<u32x4*>(&As[q][lrow][lk]) = ra[q];
      *reinterpret_cast<u32x4*>(&Bs[q][lrow][lk]) = rb[q];
    }
    __syncthreads();
    if (k0 + BK < P.K) gload(k0 + BK);
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    u32x4 a[NP][2], b[NP][2];
#pragma unroll
    for (int q = 0; q < NP; q++)
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        a[q][s2] = *reinterpret_cast<const u32x4*>(&As[q][wm * 32 + s2 * 16 + fr][fk]);
        b[q][s2] = *reinterpret_cast<const u32x4*>(&Bs[q][wn * 32 + s2 * 16 + fr][fk]);
      }
#pragma unroll
    for (int ta = 0; ta < TERMS; ta++)
#pragma unroll
      for (int tb = 0; tb < TERMS - ta; tb++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            bf16x8 ar = __builtin_bit_cast(bf16x8, a[2 * ta][i]);
            bf16x8 ai = __builtin_bit_cast(bf16x8, a[2 * ta + 1][i]);
            bf16x8 ain = __builtin_bit_cast(bf16x8, a[2 * ta + 1][i] ^ u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u});
            bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, br, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, bi, accI[i][j], 0, 0, 0);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ain, bi, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai, br, accI[i][j], 0, 0, 0);
          }
    if (TERMS >= 3) {
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
#pragma unroll
          for (int v = 0; v < 4; v++) { dR[i][j][v] += accR[i][j][v]; dI[i][j][v] += accI[i][j][v]; }
          accR[i][j] = f32x4{0, 0, 0, 0}; accI[i][j] = f32x4{0, 0, 0, 0};
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
        int n = n0 + wn * 32 + j * 16 + (lane & 15);
        epilogue_store(P, m, n, dR[i][j][v] + (double)accR[i][j][v], dI[i][j][v] + (double)accI[i][j][v]);
      }
}

int gemm_bf16planes(const GemmProb* probs_dev, int n, int total_tiles, int terms, hipStream_t s) {
  if (total_tiles <= 0) return CSX_OK;
  if (terms == 2)
    hipLaunchKernelGGL(k_gemm_planes<2>, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
  else if (terms == 3)
    hipLaunchKernelGGL(k_gemm_planes<3>, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
  else
    return fail(CSX_ERR_ARG, "gemm_bf16planes: terms must be 2 or 3");
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

__global__ void k_split_planes(const double2* __restrict__ in, unsigned short* __restrict__ out, long npix,
                               long total, int terms) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / npix, p = i - f * npix;
    double2 v = in[i];
    store_planes(out + f * 2 * terms * npix, p, npix, terms, v.x, v.y);
  }
}
int split_c128_to_planes(const void* in, void* out, long npix, int nfield, int terms, hipStream_t s) {
  long total = npix * nfield;
  if (total <= 0) return CSX_OK;
  int blocks = (int)std::min<long>(8192, (total + 255) / 256);
  hipLaunchKernelGGL(k_split_planes, dim3(blocks), dim3(256), 0, s, (const double2*)in, (unsigned short*)out, npix,
                     total, terms);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// ---------------------------------------------------------------- DFT kernels --
// M[o][i] = exp(-2 pi i t), t = (xo0 + o dxo)(xi0 + i dxi) / lf in cycles, FP64,
// reduced to [-1/2, 1/2] before sincospi so the argument keeps full accuracy.
__global__ void k_dft_matrices(const DftSpec* __restrict__ specs, int nspec) {
#pragma clang fp contract(off)
  const DftSpec& S = specs[blockIdx.y];
  long total = (long)S.nout * S.nin;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    int o = (int)(e / S.nin), i = (int)(e % S.nin);
    double t = (S.xo0 + o * S.dxo) * (S.xi0 + i * S.dxi) / S.lf;
    double f = t - rint(t);
    double sn, cs;
    sincospi(-2.0 * f, &sn, &cs);
    if (S.terms > 0) store_planes(static_cast<unsigned short*>(S.out), e, total, S.terms, cs, sn);
    else if (S.out_c64) static_cast<float2*>(S.out)[e] = make_float2((float)cs, (float)sn);
    else static_cast<double2*>(S.out)[e] = make_double2(cs, sn);
  }
}

int dft_matrices(const DftSpec* specs_dev, int n, int max_elems, hipStream_t s) {
  if (n <= 0) return CSX_OK;
  int blocks = std::min(1024, (max_elems + 255) / 256);
  hipLaunchKernelGGL(k_dft_matrices, dim3(blocks, n), dim3(256), 0, s, specs_dev, n);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

__global__ void k_c128_to_c64(const double2* __restrict__ in, float2* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
__global__ void k_c64_to_c128(const float2* __restrict__ in, double2* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = make_double2(in[i].x, in[i].y);
}
int convert_c128_to_c64(const void* in, void* out, size_t n, hipStream_t s) {
  if (!n) return CSX_OK;
  int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_c128_to_c64, dim3(blocks), dim3(256), 0, s, (const double2*)in, (float2*)out, n);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}
int convert_c64_to_c128(const void* in, void* out, size_t n, hipStream_t s) {
  if (!n) return CSX_OK;
  int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_c64_to_c128, dim3(blocks), dim3(256), 0, s, (const float2*)in, (double2*)out, n);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

}  // namespace csx
