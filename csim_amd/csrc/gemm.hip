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
        f64x2 w = gin<f64x2>(P.mask)[off];
        double tr = vr * w.x - vi * w.y;
        vi = vr * w.y + vi * w.x;
        vr = tr;
      } else {
        double w = gin<double>(P.mask)[off];
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
  if (P.out_kind == OUT_C128) gout<f64x2>(P.C)[off] = f64x2{vr, vi};
  else if (P.out_kind == OUT_C64) gout<f32x2>(P.C)[off] = f32x2{(float)vr, (float)vi};
  else store_planes(gout<unsigned short>(P.C), off, P.sC, P.terms, vr, vi);
}

// ------------------------------------------------------ FP64 / FP32 MFMA GEMM --
// Accumulator layout (measured on gfx950): value v of lane l of the FP64
// 16x16x4 MFMA is D[4v + l/16][l%16]; the FP32 16x16x4 one is D[4(l/16) + v][l%16].
template <typename T> struct Mma;
template <> struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef f64x2 c2;
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef f32x2 c2;
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
  const GLOBAL C2* A = gin<C2>(P.A);
  const GLOBAL C2* B = gin<C2>(P.B);
  const int M = P.M, N = P.N, K = P.K;
  const long lda = P.lda, ldb = P.ldb;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = t >> 2, lk = (t & 3) * 4;   // loader: 64 rows x 4 chunks of 4

  V4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) { accR[i][j] = V4{0, 0, 0, 0}; accI[i][j] = V4{0, 0, 0, 0}; }

  C2 ra[4], rb[4];
  // unconditional loads from clamped (in-bounds) addresses, zeroed by select
  const int am = m0 + lrow, bn = n0 + lrow;
  const long arow = (long)min(am, M - 1) * lda, brow = (long)min(bn, N - 1) * ldb;
  int kl = 0;   // k of the loaded slab; the select is applied at the LDS store
  auto gload = [&](int k0) {
    kl = k0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int kc = min(k0 + lk + q, K - 1);
      ra[q] = A[arow + kc];
      rb[q] = B[brow + kc];
    }
  };
  gload(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const bool kin = kl + lk + q < K;
      As[lrow][lk + q] = (am < M && kin) ? ra[q] : C2{0, 0};
      Bs[lrow][lk + q] = (bn < N && kin) ? rb[q] : C2{0, 0};
    }
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
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

// ------------------------------------------------------- split-BF16 types --
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

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
  const GLOBAL unsigned short* A = gin<unsigned short>(P.A);
  const GLOBAL unsigned short* B = gin<unsigned short>(P.B);
  const int M = P.M, N = P.N, K = P.K;
  const long sA = P.sA, sB = P.sB;

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
  // K is a multiple of 8: a chunk is wholly inside or outside.  Unconditional
  // loads from clamped addresses, zeroed by select.
  const int am = m0 + lrow, bn = n0 + lrow;
  const GLOBAL unsigned short* pa = A + (long)min(am, M - 1) * P.lda;
  const GLOBAL unsigned short* pb = B + (long)min(bn, N - 1) * P.ldb;
  bool ia = false, ib = false;   // select applied at the LDS store: loads stay in flight
  auto gload = [&](int k0) {
    const int k = k0 + lk, kc = min(k, K - 8);
    ia = am < M && k < K; ib = bn < N && k < K;
#pragma unroll
    for (int q = 0; q < NP; q++) {
      ra[q] = *reinterpret_cast<const GLOBAL u32x4*>(pa + q * sA + kc);
      rb[q] = *reinterpret_cast<const GLOBAL u32x4*>(pb + q * sB + kc);
    }
  };

  gload(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    const u32x4 z = u32x4{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < NP; q++) {
      *reinterpret_cast<u32x4*>(&As[q][lrow][lk]) = ia ? ra[q] : z;
      *reinterpret_cast<u32x4*>(&Bs[q][lrow][lk]) = ib ? rb[q] : z;
    }
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    u32x4 a[NP][2], b[NP][2];
#pragma unroll
    for (int q = 0; q < NP; q++)
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        a[q][s2] = *reinterpret_cast<const u32x4*>(&As[q][wm * 32 + s2 * 16 + fr][fk]);
        b[q][s2] = *reinterpret_cast<const u32x4*>(&Bs[q][wn * 32 + s2 * 16 + fr][fk]);
      }
    // two sweeps per product over the 2x2 sub-tiles: consecutive updates of one
    // accumulator are 8 MFMAs apart (no back-to-back dependency stalls)
    const u32x4 sgn = u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
#pragma unroll
    for (int ta = 0; ta < TERMS; ta++)
#pragma unroll
      for (int tb = 0; tb < TERMS - ta; tb++) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            bf16x8 ar = __builtin_bit_cast(bf16x8, a[2 * ta][i]);
            bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, br, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, bi, accI[i][j], 0, 0, 0);
          }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            bf16x8 ai = __builtin_bit_cast(bf16x8, a[2 * ta + 1][i]);
            bf16x8 ain = __builtin_bit_cast(bf16x8, a[2 * ta + 1][i] ^ sgn);
            bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ain, bi, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai, br, accI[i][j], 0, 0, 0);
          }
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

// 128x128 tile of the two-term (3-product) split: 4 waves of 64x64 (4x4 MFMA
// 16x16x32 sub-tiles, 32 FP32 accumulators each), K slabs of 32 double-buffered
// in LDS (2 x 64 KiB, one barrier per slab).  LDS rows are 64 B (4 chunks of
// 16 B) with chunk c of row r stored at c ^ ((r >> 1) & 3): the 8 rows a group
// of 8 lanes reads (or writes) then fall on 8 distinct 16-byte bank groups.
static_assert(kBigTileM == kBigTileN, "square big tile");
constexpr int kBigTile = kBigTileM;
__device__ __forceinline__ int swz(int row, int chunk) { return row * 32 + ((chunk ^ ((row >> 1) & 3)) << 3); }

__global__ __launch_bounds__(256, 1) void k_gemm_planes_big(const GemmProb* __restrict__ probs, int nprob) {
  constexpr int BK = 32, NP = 4, ROWS = kBigTile;
  constexpr int PLANE = ROWS * BK;                  // bf16 elements of one plane of one operand tile
  constexpr int STAGE = 2 * NP * PLANE;             // A and B, all planes
  __shared__ __align__(16) unsigned short lds[2 * STAGE];

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const GemmProb& P = probs[find_prob(probs, nprob, tile)];
  const int local = tile - P.tile_start;
  const int tm = local % P.tiles_m, tn = local / P.tiles_m;
  const int m0 = tm * kBigTile, n0 = tn * kBigTile;
  const GLOBAL unsigned short* A = gin<unsigned short>(P.A);
  const GLOBAL unsigned short* B = gin<unsigned short>(P.B);
  const int M = P.M, N = P.N, K = P.K;
  const long lda = P.lda, ldb = P.ldb, sA = P.sA, sB = P.sB;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  f32x4 accR[4][4], accI[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) { accR[i][j] = f32x4{0, 0, 0, 0}; accI[i][j] = f32x4{0, 0, 0, 0}; }

  // loader: per operand and plane 128 rows x 4 chunks = 512 chunks, 2 per thread
  u32x4 ra[NP][2], rb[NP][2];
  const GLOBAL unsigned short* pa[2];
  const GLOBAL unsigned short* pb[2];
  bool ina[2], inb[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int row = (h * 256 + t) >> 2;
    ina[h] = m0 + row < M; inb[h] = n0 + row < N;
    pa[h] = A + (long)min(m0 + row, M - 1) * lda;
    pb[h] = B + (long)min(n0 + row, N - 1) * ldb;
  }
  // K is a multiple of 8: unconditional loads from clamped addresses; the data are
  // zeroed by select only when stored to LDS, so the loads stay in flight meanwhile
  bool ka[2], kb[2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = k0 + ((h * 256 + t) & 3) * 8, kc = min(k, K - 8);
      ka[h] = ina[h] && k < K; kb[h] = inb[h] && k < K;
#pragma unroll
      for (int q = 0; q < NP; q++) {
        ra[q][h] = *reinterpret_cast<const GLOBAL u32x4*>(pa[h] + q * sA + kc);
        rb[q][h] = *reinterpret_cast<const GLOBAL u32x4*>(pb[h] + q * sB + kc);
      }
    }
  };
  auto sstore = [&](int buf) {
    unsigned short* base = lds + buf * STAGE;
    const u32x4 z = u32x4{0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int id = h * 256 + t, row = id >> 2, c = id & 3, o = swz(row, c);
#pragma unroll
      for (int q = 0; q < NP; q++) {
        *reinterpret_cast<u32x4*>(base + q * PLANE + o) = ka[h] ? ra[q][h] : z;
        *reinterpret_cast<u32x4*>(base + (NP + q) * PLANE + o) = kb[h] ? rb[q][h] : z;
      }
    }
  };

  const int nslab = (K + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int fr = lane & 15, fc = lane >> 4;
  const u32x4 sign = u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  for (int s = 0; s < nslab; s++) {
    const int cur = s & 1;
    if (s + 1 < nslab) gload((s + 1) * BK);
    const unsigned short* base = lds + cur * STAGE;
    u32x4 b[NP][4];
#pragma unroll
    for (int q = 0; q < NP; q++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        b[q][j] = *reinterpret_cast<const u32x4*>(base + (NP + q) * PLANE + swz(wn * 64 + j * 16 + fr, fc));
#pragma unroll
    for (int i = 0; i < 4; i++) {
      u32x4 a[NP];
#pragma unroll
      for (int q = 0; q < NP; q++)
        a[q] = *reinterpret_cast<const u32x4*>(base + q * PLANE + swz(wm * 64 + i * 16 + fr, fc));
#pragma unroll
      // two sweeps per product over j: consecutive updates of one accumulator are
      // 8 MFMAs apart (no back-to-back dependency stalls with one wave per SIMD)
      for (int ta = 0; ta < 2; ta++) {
        const bf16x8 ar = __builtin_bit_cast(bf16x8, a[2 * ta]);
        const bf16x8 ai = __builtin_bit_cast(bf16x8, a[2 * ta + 1]);
        const bf16x8 ain = __builtin_bit_cast(bf16x8, a[2 * ta + 1] ^ sign);
#pragma unroll
        for (int tb = 0; tb < 2 - ta; tb++) {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            const bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, br, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, bi, accI[i][j], 0, 0, 0);
          }
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            const bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ain, bi, accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai, br, accI[i][j], 0, 0, 0);
          }
        }
      }
    }
    if (s + 1 < nslab) sstore(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int m = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + v;
        int n = n0 + wn * 64 + j * 16 + (lane & 15);
        epilogue_store(P, m, n, (double)accR[i][j][v], (double)accI[i][j][v]);
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

int gemm_bf16planes_big(const GemmProb* probs_dev, int n, int total_tiles, hipStream_t s) {
  if (total_tiles <= 0) return CSX_OK;
  hipLaunchKernelGGL(k_gemm_planes_big, dim3(total_tiles), dim3(256), 0, s, probs_dev, n);
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
