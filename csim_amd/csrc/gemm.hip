// Grouped complex GEMM kernels on CDNA4 (gfx950) MFMA and DFT-kernel generation.
//
// Every matrix Fourier transform of the path (csim zoomFft, src/lib/csim_fft.cpp:
// 670-826, which csim evaluates with a Bluestein chirp-z transform) is computed in
// its closed form as two complex GEMMs.  One launch covers every field of a batch
// (wavelength x source x candidate, uneven sizes allowed): a grouped GEMM whose
// workgroups find their problem by binary search over per-problem tile offsets.
#include "gemm.hpp"
#include "planes.hpp"

#include <hip/hip_bf16.h>
#include <algorithm>
#include <vector>

namespace csx {

// ------------------------------------------------------------ problem plan --
int plan_tiles(GemmProb* probs, int n, int tile_m, int tile_n) {
  int total = 0;
  long elems = 0;
  for (int i = 0; i < n; i++) {
    probs[i].tiles_m = (probs[i].M + tile_m - 1) / tile_m;
    probs[i].tiles_n = (probs[i].N + tile_n - 1) / tile_n;
    probs[i].tile_start = total;
    probs[i].elem_start = elems;
    total += probs[i].tiles_m * probs[i].tiles_n * (probs[i].ksplit > 1 ? probs[i].ksplit : 1);
    elems += (long)probs[i].M * probs[i].N;
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

// K range of a BM x BN tile from the problem's block table (GemmProb::kr): whole
// slabs of BK from a multiple of BK, so the slab grid, and with it the order of
// every sum, stays that of the full range; only products that are exact zeros drop.
// Returns true when every block of the tile is zero and never read (lo = -1): the
// tile writes nothing.
template <int BM, int BN, int BK>
__device__ __forceinline__ bool tile_krange(const GemmProb& P, int m0, int n0, int& kbeg, int& kend) {
  if (!P.kr) return false;
  const GLOBAL int* kr = gin<int>(P.kr);
  int lo = P.K, hi = 0;
  bool skip = true;
  const int mb1 = min((m0 + BM + kKrBlock - 1) / kKrBlock, (P.M + kKrBlock - 1) / kKrBlock);
  const int nb1 = min((n0 + BN + kKrBlock - 1) / kKrBlock, P.kr_ld);
  for (int mb = m0 / kKrBlock; mb < mb1; mb++)
    for (int nb = n0 / kKrBlock; nb < nb1; nb++) {
      const long b = 2 * ((long)mb * P.kr_ld + nb);
      skip = skip && kr[b] < 0;
      lo = min(lo, kr[b] < 0 ? P.K : kr[b]);
      hi = max(hi, kr[b + 1]);
    }
  kbeg = max(kbeg, lo / BK * BK);
  kend = max(kbeg, min(kend, hi));
  return skip;
}

// K range of split ks of a problem split ksplit ways: ceil(K / ksplit) rounded up to
// whole slabs of BK, so every split but the last starts and ends on the slab grid.
template <int BK>
__device__ __forceinline__ void split_krange(const GemmProb& P, int ks, int& kbeg, int& kend) {
  kbeg = 0;
  kend = P.K;
  if (P.ksplit > 1) {
    const int kc = ((P.K + P.ksplit - 1) / P.ksplit + BK - 1) / BK * BK;
    kbeg = min(P.K, ks * kc);
    kend = min(P.K, kbeg + kc);
  }
}
// One split's FP64 sums into its workspace slice (the finishing launch adds them).
__device__ __forceinline__ void split_store(const GemmProb& P, int ks, int m, int n, double re, double im) {
  if (m < P.M && n < P.N)
    gout<f64x2>(P.ws)[((long)ks * P.M + m) * P.N + n] = f64x2{re, im};
}

// Epilogue: alpha, mask (epilogue_value), store (shared by every kernel).
__device__ __forceinline__ void epilogue_value(const GemmProb& P, int m, int n, double re, double im, double& out_r,
                                               double& out_i) {
  double vr = P.alpha_re * re - P.alpha_im * im;
  double vi = P.alpha_re * im + P.alpha_im * re;
  if (P.addend) {
    const long off = P.transC ? (long)n * P.ldc + m : (long)m * P.ldc + n;
    if (P.out_kind == OUT_C128) {
      const f64x2 a = gin<f64x2>(P.addend)[off];
      vr += a.x; vi += a.y;
    } else if (P.out_kind == OUT_C64) {
      const f32x2 a = gin<f32x2>(P.addend)[off];
      vr += a.x; vi += a.y;
    } else {
      const GLOBAL unsigned short* a = gin<unsigned short>(P.addend);
      double ar = 0.0, ai = 0.0;
      for (int t = 0; t < P.terms; t++) {
        ar += (double)bf16_val(a[(2 * t) * P.sC + off]);
        ai += (double)bf16_val(a[(2 * t + 1) * P.sC + off]);
      }
      vr += ar; vi += ai;
    }
  }
  if (P.mask_kind != MASK_NONE) {
    int mr = m - P.mm0, nc = n - P.mn0;
    if (mr >= 0 && mr < P.mM && nc >= 0 && nc < P.mN) {
      long off = (long)mr * P.ldm + nc;
      if (P.mask_kind == MASK_COMPLEX || P.mask_kind == MASK_COMPLEX_CLIP) {
        f64x2 w = gin<f64x2>(P.mask)[off];
        double tr = vr * w.x - vi * w.y;
        vi = vr * w.y + vi * w.x;
        vr = tr;
      } else {
        const double w = P.mask_kind == MASK_REAL_1PI_U8    ? (double)gin<unsigned char>(P.mask)[off]
                         : P.mask_kind == MASK_REAL_1PI_F32 ? (double)gin<float>(P.mask)[off]
                                                            : gin<double>(P.mask)[off];
        if (P.mask_kind != MASK_REAL) {   // csim mask: (1+i) w
          double tr = (vr - vi) * w;
          vi = (vr + vi) * w;
          vr = tr;
        } else {
          vr *= w; vi *= w;
        }
      }
    } else if (P.mask_kind == MASK_COMPLEX_CLIP) {
      vr = 0; vi = 0;
    }
  }
  out_r = vr;
  out_i = vi;
}
// zero: the tile took no K, which the K range tables give only where the products are
// exact zeros and no addend is bound, or where the epilogue mask is zero: the output is
// zero and the mask is not read
__device__ __forceinline__ void epilogue_store(const GemmProb& P, int m, int n, double re, double im,
                                               bool zero = false) {
  if (m >= P.M || n >= P.N) return;
  double vr = 0.0, vi = 0.0;
  if (!zero) epilogue_value(P, m, n, re, im, vr, vi);
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
  // by value: the descriptor's fields then live in scalar registers, not reloaded after
  // every store (a store through one of its pointers could alias a reference to it)
  const GemmProb P = probs[find_prob(probs, nprob, tile)];
  const int local = tile - P.tile_start, tiles_mn = P.tiles_m * P.tiles_n;
  const int ks = local / tiles_mn, lmn = local - ks * tiles_mn;   // K split, output tile
  const int tm = lmn % P.tiles_m, tn = lmn / P.tiles_m;
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

  int kbeg, kend;
  split_krange<BK>(P, ks, kbeg, kend);
  if (tile_krange<kTileM, kTileN, BK>(P, m0, n0, kbeg, kend) && P.ksplit > 1) return;
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
  gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const bool kin = kl + lk + q < K;
      As[lrow][lk + q] = (am < M && kin) ? ra[q] : C2{0, 0};
      Bs[lrow][lk + q] = (bn < N && kin) ? rb[q] : C2{0, 0};
    }
    __syncthreads();
    if (k0 + BK < kend) gload(k0 + BK);
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
        if (P.ksplit > 1) split_store(P, ks, m, n, (double)accR[i][j][v], (double)accI[i][j][v]);
        else epilogue_store(P, m, n, (double)accR[i][j][v], (double)accI[i][j][v], kend <= kbeg);
      }
}

// FP64 tile variants: block BM x BN, WM x WN waves of IM x JN 16x16 sub-tiles, K
// slabs of 16 complex values double-buffered in LDS (one barrier per slab, the next
// slab's global loads in flight during the current slab's MFMAs).  Within a 4-deep k
// step the products are issued in two sweeps (real parts of A, then imaginary
// parts), so two updates of one accumulator are 2 IM JN MFMAs apart.
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_f64db(const GemmProb* __restrict__ probs, int nprob) {
  typedef Mma<double>::v4 V4;
  constexpr int NT = 64 * WM * WN, BK = 16, LDK = BK + 1;
  constexpr int IM = BM / WM / 16, JN = BN / WN / 16;
  constexpr int CA = BM * BK / 4 / NT, CB = BN * BK / 4 / NT;   // 4-value chunks per thread
  static_assert(IM >= 1 && JN >= 1 && CA >= 1 && CB >= 1 && (BM * BK / 4) % NT == 0 && (BN * BK / 4) % NT == 0,
                "tile shape");
  __shared__ f64x2 As[2][BM][LDK];
  __shared__ f64x2 Bs[2][BN][LDK];

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  // by value: the descriptor's fields then live in scalar registers, not reloaded after
  // every store (a store through one of its pointers could alias a reference to it)
  const GemmProb P = probs[find_prob(probs, nprob, tile)];
  const int local = tile - P.tile_start, tiles_mn = P.tiles_m * P.tiles_n;
  const int ks = local / tiles_mn, lmn = local - ks * tiles_mn;   // K split, output tile
  const int tm = lmn % P.tiles_m, tn = lmn / P.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const GLOBAL f64x2* A = gin<f64x2>(P.A);
  const GLOBAL f64x2* B = gin<f64x2>(P.B);
  const int M = P.M, N = P.N, K = P.K;
  const long lda = P.lda, ldb = P.ldb;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN;

  V4 accR[IM][JN], accI[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; i++)
#pragma unroll
    for (int j = 0; j < JN; j++) { accR[i][j] = V4{0, 0, 0, 0}; accI[i][j] = V4{0, 0, 0, 0}; }

  // chunk e = c * NT + t: row e / 4, k offset (e % 4) * 4 (4 threads cover a 16-deep row)
  f64x2 ra[CA][4], rb[CB][4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < CA; c++) {
      const int e = c * NT + t, r = e >> 2, kq = (e & 3) * 4;
      const long row = (long)min(m0 + r, M - 1) * lda;
#pragma unroll
      for (int q = 0; q < 4; q++) ra[c][q] = A[row + min(k0 + kq + q, K - 1)];
    }
#pragma unroll
    for (int c = 0; c < CB; c++) {
      const int e = c * NT + t, r = e >> 2, kq = (e & 3) * 4;
      const long row = (long)min(n0 + r, N - 1) * ldb;
#pragma unroll
      for (int q = 0; q < 4; q++) rb[c][q] = B[row + min(k0 + kq + q, K - 1)];
    }
  };
  // the zeroing of rows / k beyond the matrix is applied when the slab is stored
  auto sstore = [&](int buf, int k0) {
    const f64x2 z = f64x2{0.0, 0.0};
#pragma unroll
    for (int c = 0; c < CA; c++) {
      const int e = c * NT + t, r = e >> 2, kq = (e & 3) * 4;
      const bool rin = m0 + r < M;
#pragma unroll
      for (int q = 0; q < 4; q++) As[buf][r][kq + q] = (rin && k0 + kq + q < K) ? ra[c][q] : z;
    }
#pragma unroll
    for (int c = 0; c < CB; c++) {
      const int e = c * NT + t, r = e >> 2, kq = (e & 3) * 4;
      const bool rin = n0 + r < N;
#pragma unroll
      for (int q = 0; q < 4; q++) Bs[buf][r][kq + q] = (rin && k0 + kq + q < K) ? rb[c][q] : z;
    }
  };
  int kbeg, kend;
  split_krange<BK>(P, ks, kbeg, kend);
  if (tile_krange<BM, BN, BK>(P, m0, n0, kbeg, kend) && P.ksplit > 1) return;
  const int nslab = (kend - kbeg + BK - 1) / BK;
  gload(kbeg);
  sstore(0, kbeg);
  __syncthreads();
  for (int sl = 0; sl < nslab; sl++) {
    const int cur = sl & 1;
    if (sl + 1 < nslab) gload(kbeg + (sl + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      f64x2 a[IM], b[JN];
#pragma unroll
      for (int i = 0; i < IM; i++) a[i] = As[cur][wm * (BM / WM) + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < JN; j++) b[j] = Bs[cur][wn * (BN / WN) + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < IM; i++)
#pragma unroll
        for (int j = 0; j < JN; j++) {
          accR[i][j] = Mma<double>::mma(a[i].x, b[j].x, accR[i][j]);
          accI[i][j] = Mma<double>::mma(a[i].x, b[j].y, accI[i][j]);
        }
#pragma unroll
      for (int i = 0; i < IM; i++)
#pragma unroll
        for (int j = 0; j < JN; j++) {
          accR[i][j] = Mma<double>::mma(-a[i].y, b[j].y, accR[i][j]);
          accI[i][j] = Mma<double>::mma(a[i].y, b[j].x, accI[i][j]);
        }
    }
    if (sl + 1 < nslab) sstore(cur ^ 1, kbeg + (sl + 1) * BK);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < IM; i++)
#pragma unroll
    for (int j = 0; j < JN; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = m0 + wm * (BM / WM) + i * 16 + Mma<double>::row(lane, v);
        const int n = n0 + wn * (BN / WN) + j * 16 + (lane & 15);
        if (P.ksplit > 1) split_store(P, ks, m, n, accR[i][j][v], accI[i][j][v]);
        else epilogue_store(P, m, n, accR[i][j][v], accI[i][j][v], nslab == 0);
      }
}

struct F64Variant {
  int bm, bn, threads;
  void (*kernel)(const GemmProb*, int);
  const char* name;
};
static const F64Variant kVariantsF64[] = {
    {64, 64, 256, k_gemm_nt<double>, "nt"},
    {64, 64, 256, k_gemm_f64db<64, 64, 2, 2>, "64,64,2,2"},
    {64, 128, 256, k_gemm_f64db<64, 128, 2, 2>, "64,128,2,2"},
    {128, 64, 256, k_gemm_f64db<128, 64, 2, 2>, "128,64,2,2"},
    {128, 128, 512, k_gemm_f64db<128, 128, 2, 4>, "128,128,2,4"},
};
int f64_variant_count() { return (int)(sizeof(kVariantsF64) / sizeof(kVariantsF64[0])); }
int f64_variant_tile(int v, int* bm, int* bn) {
  if (v < 0 || v >= f64_variant_count()) return fail(CSX_ERR_ARG, "fp64 variant out of range");
  *bm = kVariantsF64[v].bm;
  *bn = kVariantsF64[v].bn;
  return CSX_OK;
}

int gemm_f64(const GemmProb* probs_dev, int n, int total_tiles, hipStream_t s, int variant) {
  if (total_tiles <= 0) return CSX_OK;
  if (variant < 0 || variant >= f64_variant_count()) return fail(CSX_ERR_ARG, "gemm_f64: bad variant");
  const F64Variant& v = kVariantsF64[variant];
  hipLaunchKernelGGL(v.kernel, dim3(total_tiles), dim3(v.threads), 0, s, probs_dev, n);
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
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// ------------------------------------------------- BF16-plane MFMA GEMM --
// Both operands arrive as BF16 planes (split once: DFT kernels at plan time,
// fields by the producing epilogue), so the main loop is 16-byte copies into
// LDS and v_mfma_f32_16x16x32_bf16; products kept: (a, b) terms with a + b < TERMS.
// FP32 accumulators are flushed into FP64 after every K slab of 32 when TERMS == 3.
//
// One template covers the tile variants the tuner chooses from: block tile
// BM x BN, WM x WN waves (each (BM/WM) x (BN/WN) = IM x JN MFMA sub-tiles),
// single (one barrier pair per slab, more blocks per CU) or double-buffered LDS
// (one barrier per slab).  LDS rows are 64 B (4 chunks of 16 B) with chunk c of
// row r stored at c ^ ((r >> 1) & 3): the 8 rows a group of 8 lanes reads or
// writes fall on 8 distinct 16-byte bank groups.
// With BK = 64 the rows are 128 B (8 chunks) and chunk c of row r sits at c ^ (r & 7).
template <int BK>
__device__ __forceinline__ int swz(int row, int chunk) {
  return BK == 32 ? row * 32 + ((chunk ^ ((row >> 1) & 3)) << 3) : row * 64 + ((chunk ^ (row & 7)) << 3);
}

// BF16 MFMA shapes (layouts measured on gfx950).  16x16x32: lane l holds row l%16,
// k = 8 (l/16) .. +7 of each operand; D value v at [4 (l/16) + v][l%16].
// 32x32x16: lane l holds row l%32, k = 8 (l/32) .. +7; D value v at
// [4 (l/32) + (v%4) + 8 (v/4)][l%32].
typedef float f32x16 __attribute__((ext_vector_type(16)));
template <int MF> struct MfmaBf16;
template <> struct MfmaBf16<16> {
  typedef f32x4 acc_t;
  static constexpr int NV = 4, KSTEP = 32;
  static __device__ __forceinline__ acc_t mma(bf16x8 a, bf16x8 b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int frow(int lane) { return lane & 15; }
  static __device__ __forceinline__ int fchunk(int lane, int ks) { return ks * 4 + (lane >> 4); }
  static __device__ __forceinline__ int drow(int lane, int v) { return 4 * (lane >> 4) + v; }
  static __device__ __forceinline__ int dcol(int lane) { return lane & 15; }
};
template <> struct MfmaBf16<32> {
  typedef f32x16 acc_t;
  static constexpr int NV = 16, KSTEP = 16;
  static __device__ __forceinline__ acc_t mma(bf16x8 a, bf16x8 b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int frow(int lane) { return lane & 31; }
  static __device__ __forceinline__ int fchunk(int lane, int ks) { return ks * 2 + (lane >> 5); }
  static __device__ __forceinline__ int drow(int lane, int v) { return 4 * (lane >> 5) + (v & 3) + 8 * (v >> 2); }
  static __device__ __forceinline__ int dcol(int lane) { return lane & 31; }
};

// Position of output column nl of tile row ml in the LDS staging of plane epilogues
// (rows of BN >= 16 elements, BN a multiple of 16; a permutation of each row's aligned
// 16-element groups).  The finishing loop has thread c of a row read columns 8 c + k,
// k = 0..7: the chunk index XORed into the low bits spreads those 8 lanes over 8
// distinct positions of a 16-element group for every k, and the row parity in bit 3
// keeps the next row's 8 lanes off them (the 16 lanes a 128-byte LDS cycle serves for
// 8-byte elements, or the 8 lanes of one row for 16-byte ones).  A store instruction of
// the MFMA layout writes 16 consecutive columns of one row: they stay distinct.
__device__ __forceinline__ int stage_col(int ml, int nl) { return nl ^ ((nl >> 3) & 7) ^ ((ml & 1) << 3); }

// Zero chunk that direct-to-LDS loads of k beyond K read instead of the matrix.
static __device__ __align__(16) unsigned int kZeroChunk[4];

// s_waitcnt with vmcnt(N) (expcnt and lgkmcnt not waited for): at most N vector
// memory instructions of this wave still in flight.
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (((N >> 4) & 3) << 14) | 0x0F70);
}

// DMA: 0 loads through registers; 1 direct-to-LDS loads into a two-slot ring (the next
// slab in flight during a slab's MFMAs); R >= 3 an R-slot ring, R - 1 slabs in flight.
template <int TERMS, int BM, int BN, int WM, int WN, bool DB, int WPE, int BK, int MF, int DMA = 0>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_gemm_planes(const GemmProb* __restrict__ probs, int nprob) {
  typedef MfmaBf16<MF> Mf;
  typedef typename Mf::acc_t Acc;
  constexpr int NV = Mf::NV;
  constexpr int NT = 64 * WM * WN, NP = 2 * TERMS, CH = BK / 8;   // 16-byte chunks per LDS row
  constexpr int IM = BM / WM / MF, JN = BN / WN / MF;
  static_assert(IM >= 1 && JN >= 1 && (BM / WM) % MF == 0 && (BN / WN) % MF == 0, "wave tile");
  constexpr int PA = BM * BK, PB = BN * BK;        // bf16 elements of one plane of each operand tile
  constexpr int STAGE = NP * (PA + PB);
  constexpr int ITA = NP * BM * CH / NT, ITB = NP * BN * CH / NT;
  static_assert(BK == 32 || BK == 64, "slab depth");
  static_assert((NP * BM * CH) % NT == 0 && (NP * BN * CH) % NT == 0 && NT % CH == 0, "loader shape");
  static_assert(TERMS == 2 || IM * JN * Mf::NV <= 32, "FP64 flush registers");
  constexpr int RING = DMA == 0 ? 0 : (DMA < 2 ? 2 : DMA);
  static_assert(RING <= 4, "ring depth");
  constexpr int NBUF = DMA ? RING : (DB ? 2 : 1);   // LDS stages
  __shared__ __align__(16) unsigned short lds[NBUF * STAGE];

  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  // by value: the descriptor's fields then live in scalar registers, not reloaded after
  // every store (a store through one of its pointers could alias a reference to it)
  const GemmProb P = probs[find_prob(probs, nprob, tile)];
  const int tiles_mn = P.tiles_m * P.tiles_n;
  const int local = tile - P.tile_start;
  const int ks = local / tiles_mn, lmn = local - ks * tiles_mn;   // K split, output tile
  const int tm = lmn % P.tiles_m, tn = lmn / P.tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  const GLOBAL unsigned short* A = gin<unsigned short>(P.A);
  const GLOBAL unsigned short* B = gin<unsigned short>(P.B);
  const int M = P.M, N = P.N, K = P.K;
  // K range of this block: whole slabs except at K itself
  int kbeg, kend;
  split_krange<BK>(P, ks, kbeg, kend);
  if (tile_krange<BM, BN, BK>(P, m0, n0, kbeg, kend)) return;   // uniform over the block

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN;

  // loader: chunk e = it * NT + t of the (plane, row, chunk) list of each operand
  // (the chunk index c = t % CH is the same for every it since NT % CH == 0);
  // element offsets within the field (< 2^31) are fixed per it
  const int ch = (t % CH) * 8;
  unsigned offA[ITA], offB[ITB];
  unsigned okA = 0, okB = 0;                      // row-in-range bits
#pragma unroll
  for (int it = 0; it < ITA; it++) {
    const int e = it * NT + t, q = e / (BM * CH), r = (e % (BM * CH)) / CH;
    if (m0 + r < M) okA |= 1u << it;
    offA[it] = (unsigned)(q * P.sA + (long)min(m0 + r, M - 1) * P.lda + ch);
  }
#pragma unroll
  for (int it = 0; it < ITB; it++) {
    const int e = it * NT + t, q = e / (BN * CH), r = (e % (BN * CH)) / CH;
    if (n0 + r < N) okB |= 1u << it;
    offB[it] = (unsigned)(q * P.sB + (long)min(n0 + r, N - 1) * P.ldb + ch);
  }
  u32x4 ra[ITA], rb[ITB];
  unsigned vA = 0, vB = 0;                        // valid bits of the slab in registers
  const bool tile_full = m0 + BM <= M && n0 + BN <= N;
  bool full = false;                              // block-uniform: no chunk of the slab needs zeroing
  // K is a multiple of 8: a chunk is wholly inside or outside.  Loads are
  // unconditional (clamped addresses); the data are zeroed only when stored to
  // LDS, so the loads stay in flight meanwhile.
  auto gload = [&](int k0) {
    full = tile_full && k0 + BK <= K;
    if (k0 + BK <= K) {
#pragma unroll
      for (int it = 0; it < ITA; it++) ra[it] = *reinterpret_cast<const GLOBAL u32x4*>(A + offA[it] + k0);
#pragma unroll
      for (int it = 0; it < ITB; it++) rb[it] = *reinterpret_cast<const GLOBAL u32x4*>(B + offB[it] + k0);
      vA = okA; vB = okB;
    } else {
      const int k = k0 + ch, kc = min(k, K - 8) - ch;
      const bool kin = k < K;
#pragma unroll
      for (int it = 0; it < ITA; it++) ra[it] = *reinterpret_cast<const GLOBAL u32x4*>(A + offA[it] + kc);
#pragma unroll
      for (int it = 0; it < ITB; it++) rb[it] = *reinterpret_cast<const GLOBAL u32x4*>(B + offB[it] + kc);
      vA = kin ? okA : 0u; vB = kin ? okB : 0u;
    }
  };
  auto sstore = [&](int buf) {
    unsigned short* base = lds + buf * STAGE;
    const u32x4 z = u32x4{0, 0, 0, 0};
    const int c = t % CH;
    if (full) {
#pragma unroll
      for (int it = 0; it < ITA; it++) {
        const int e = it * NT + t, q = e / (BM * CH), r = (e % (BM * CH)) / CH;
        *reinterpret_cast<u32x4*>(base + q * PA + swz<BK>(r, c)) = ra[it];
      }
#pragma unroll
      for (int it = 0; it < ITB; it++) {
        const int e = it * NT + t, q = e / (BN * CH), r = (e % (BN * CH)) / CH;
        *reinterpret_cast<u32x4*>(base + NP * PA + q * PB + swz<BK>(r, c)) = rb[it];
      }
      return;
    }
#pragma unroll
    for (int it = 0; it < ITA; it++) {
      const int e = it * NT + t, q = e / (BM * CH), r = (e % (BM * CH)) / CH;
      *reinterpret_cast<u32x4*>(base + q * PA + swz<BK>(r, c)) = (vA >> it) & 1 ? ra[it] : z;
    }
#pragma unroll
    for (int it = 0; it < ITB; it++) {
      const int e = it * NT + t, q = e / (BN * CH), r = (e % (BN * CH)) / CH;
      *reinterpret_cast<u32x4*>(base + NP * PA + q * PB + swz<BK>(r, c)) = (vB >> it) & 1 ? rb[it] : z;
    }
  };

  // Direct-to-LDS loads (global_load_lds_dwordx4): instruction i of wave w fills the
  // 64 consecutive 16-byte LDS chunks (i NW + w) * 64 + lane of the stage, so each lane
  // fetches the global chunk that belongs in its slot (the swizzle inverted: slot
  // (row r, position p) holds chunk p ^ ((r >> 1) & 3)).  Rows beyond the matrix
  // read a clamped row (those outputs are never stored); k beyond K reads zeros.
  constexpr int NW = WM * WN, CA = NP * BM * CH, CB = NP * BN * CH, NI = DMA ? (CA + CB) / NT : 1;
  static_assert(!DMA || (DB && BK == 32 && CA % (NW * 64) == 0 && CB % (NW * 64) == 0), "direct-load shape");
  unsigned doff[NI];
  int dk[NI];
  if (DMA) {
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int sl = (i * NW + wave) * 64 + lane;
      const bool isA = sl < CA;
      const int e = isA ? sl : sl - CA, rows = isA ? BM : BN;
      const int q = e / (rows * CH), rem = e % (rows * CH), r = rem / CH, c = (rem % CH) ^ ((r >> 1) & 3);
      dk[i] = c * 8;
      doff[i] = isA ? (unsigned)(q * P.sA + (long)min(m0 + r, M - 1) * P.lda + c * 8)
                    : (unsigned)(q * P.sB + (long)min(n0 + r, N - 1) * P.ldb + c * 8);
    }
  }
  auto dload = [&](int k0, int buf) {
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const bool isA = (i * NW) * 64 < CA;              // uniform: whole instructions per operand
      const GLOBAL unsigned short* src = (isA ? A : B) + doff[i] + k0;
      const GLOBAL void* g = k0 + dk[i] < K ? (const GLOBAL void*)src : (const GLOBAL void*)kZeroChunk;
      __attribute__((address_space(3))) void* l =
          (__attribute__((address_space(3))) void*)(lds + buf * STAGE + (i * NW + wave) * 64 * 8);
      __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
    }
  };

  // Three terms: the products of the two leading (fixed-point) parts go to their own
  // accumulators hR/hI, whose sums over one K step are exact (planes.hpp); both sets
  // are added into the FP64 sums dR/dI after every K step.
  Acc accR[IM][JN], accI[IM][JN], hR[IM][JN], hI[IM][JN];
  double dR[TERMS == 3 ? IM : 1][TERMS == 3 ? JN : 1][TERMS == 3 ? NV : 1];
  double dI[TERMS == 3 ? IM : 1][TERMS == 3 ? JN : 1][TERMS == 3 ? NV : 1];
#pragma unroll
  for (int i = 0; i < IM; i++)
#pragma unroll
    for (int j = 0; j < JN; j++)
#pragma unroll
      for (int v = 0; v < NV; v++) { accR[i][j][v] = 0.f; accI[i][j][v] = 0.f; hR[i][j][v] = 0.f; hI[i][j][v] = 0.f; }
  if (TERMS == 3) {
#pragma unroll
    for (int i = 0; i < (TERMS == 3 ? IM : 1); i++)
#pragma unroll
      for (int j = 0; j < (TERMS == 3 ? JN : 1); j++)
#pragma unroll
        for (int v = 0; v < (TERMS == 3 ? NV : 1); v++) { dR[i][j][v] = 0; dI[i][j][v] = 0; }
  }

  const int fr = Mf::frow(lane);
  const u32x4 sgn = u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  auto compute = [&](const unsigned short* base) {
#pragma unroll
   for (int kk = 0; kk < BK / Mf::KSTEP; kk++) {
    const int fk = Mf::fchunk(lane, kk);         // 16-byte chunk of this lane's 8 k values
    u32x4 b[NP][JN];
#pragma unroll
    for (int q = 0; q < NP; q++)
#pragma unroll
      for (int j = 0; j < JN; j++)
        b[q][j] = *reinterpret_cast<const u32x4*>(base + NP * PA + q * PB + swz<BK>(wn * (BN / WN) + j * MF + fr, fk));
#pragma unroll
    for (int i = 0; i < IM; i++) {
      u32x4 a[NP];
#pragma unroll
      for (int q = 0; q < NP; q++) a[q] = *reinterpret_cast<const u32x4*>(base + q * PA + swz<BK>(wm * (BM / WM) + i * MF + fr, fk));
#pragma unroll
      for (int ta = 0; ta < TERMS; ta++) {
        const bf16x8 ar = __builtin_bit_cast(bf16x8, a[2 * ta]);
        const bf16x8 ai = __builtin_bit_cast(bf16x8, a[2 * ta + 1]);
        const bf16x8 ain = __builtin_bit_cast(bf16x8, a[2 * ta + 1] ^ sgn);
#pragma unroll
        for (int tb = 0; tb < TERMS - ta; tb++) {
          Acc (&xR)[IM][JN] = (TERMS == 3 && ta == 0 && tb == 0) ? hR : accR;
          Acc (&xI)[IM][JN] = (TERMS == 3 && ta == 0 && tb == 0) ? hI : accI;
          // two sweeps over j: updates of one accumulator are 2 JN MFMAs apart
#pragma unroll
          for (int j = 0; j < JN; j++) {
            const bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            const bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            xR[i][j] = Mf::mma(ar, br, xR[i][j]);
            xI[i][j] = Mf::mma(ar, bi, xI[i][j]);
          }
#pragma unroll
          for (int j = 0; j < JN; j++) {
            const bf16x8 br = __builtin_bit_cast(bf16x8, b[2 * tb][j]);
            const bf16x8 bi = __builtin_bit_cast(bf16x8, b[2 * tb + 1][j]);
            xR[i][j] = Mf::mma(ain, bi, xR[i][j]);
            xI[i][j] = Mf::mma(ai, br, xI[i][j]);
          }
        }
      }
    }
    // the leading products' sums are exact within one K window only: flushed after each
    if (TERMS == 3 && ((kk + 1) * Mf::KSTEP) % kPlaneWindow == 0) {
#pragma unroll
      for (int i = 0; i < (TERMS == 3 ? IM : 1); i++)
#pragma unroll
        for (int j = 0; j < (TERMS == 3 ? JN : 1); j++) {
#pragma unroll
          for (int v = 0; v < (TERMS == 3 ? NV : 1); v++) {
            dR[i][j][v] += (double)hR[i][j][v];
            dI[i][j][v] += (double)hI[i][j][v];
            hR[i][j][v] = 0.f; hI[i][j][v] = 0.f;
          }
        }
    }
   }
  };
  // The other products are 2^-8 of the leading ones (at most 2^-16 for the terms kept),
  // so their FP32 sums over 4 slabs (128 values) round at about 2^-30 of the leading
  // sum, below the 2^-24 of the products left out: they are flushed into FP64 every 4
  // slabs and at the end.
  auto flush_rest = [&]() {
    if (TERMS == 3) {
#pragma unroll
      for (int i = 0; i < (TERMS == 3 ? IM : 1); i++)
#pragma unroll
        for (int j = 0; j < (TERMS == 3 ? JN : 1); j++) {
#pragma unroll
          for (int v = 0; v < (TERMS == 3 ? NV : 1); v++) {
            dR[i][j][v] += (double)accR[i][j][v];
            dI[i][j][v] += (double)accI[i][j][v];
            accR[i][j][v] = 0.f; accI[i][j][v] = 0.f;
          }
        }
    }
  };
  constexpr int kRestSlabs = 4;

  const int nslab = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (nslab > 0) {
  if (DMA) {
    // slab s lives in slot s % RING; before slab s is read every wave waits until at
    // most the loads of the slabs issued after it are in flight (NI instructions each)
    auto wait_after = [&](int later) {
      if (RING >= 4 && later >= 2) wait_vmcnt<(RING >= 4 ? 2 * NI : 0)>();
      else if (RING >= 3 && later >= 1) wait_vmcnt<(RING >= 3 ? NI : 0)>();
      else wait_vmcnt<0>();
    };
    for (int q = 0; q < RING - 1 && q < nslab; q++) dload(kbeg + q * BK, q);
    wait_after(min(RING - 1, nslab) - 1);
    __syncthreads();
    for (int s = 0; s < nslab; s++) {
      const int nx = s + RING - 1;                 // slot (s - 1) % RING: read by every wave before the last barrier
      if (nx < nslab) dload(kbeg + nx * BK, nx % RING);
      compute(lds + (s % RING) * STAGE);
      if (s % kRestSlabs == kRestSlabs - 1) flush_rest();
      if (s + 1 < nslab) wait_after(min(nx, nslab - 1) - (s + 1));
      __syncthreads();
    }
  } else if (DB) {
    gload(kbeg);
    sstore(0);
    __syncthreads();
    for (int s = 0; s < nslab; s++) {
      const int cur = s & 1;
      if (s + 1 < nslab) gload(kbeg + (s + 1) * BK);
      compute(lds + cur * STAGE);
      if (s % kRestSlabs == kRestSlabs - 1) flush_rest();
      if (s + 1 < nslab) sstore(cur ^ 1);
      __syncthreads();
    }
  } else {
    gload(kbeg);
    for (int s = 0; s < nslab; s++) {
      __syncthreads();
      sstore(0);
      __syncthreads();
      if (s + 1 < nslab) gload(kbeg + (s + 1) * BK);
      compute(lds);
      if (s % kRestSlabs == kRestSlabs - 1) flush_rest();
    }
  }
  }
  flush_rest();
  if (TERMS == 3 && P.ksplit > 1) {
    // this split's FP64 sums; gemm_splitk_finish adds the splits and applies the epilogue
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < NV; v++)
          split_store(P, ks, m0 + wm * (BM / WM) + i * MF + Mf::drow(lane, v),
                      n0 + wn * (BN / WN) + j * MF + Mf::dcol(lane),
                      (double)accR[i][j][v] + dR[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v],
                      (double)accI[i][j][v] + dI[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v]);
    return;
  }
  if (TERMS == 2 && P.ksplit > 1) {
    // this split's FP32 sums; gemm_splitk_finish adds the splits and applies the epilogue
    float2* W = reinterpret_cast<float2*>(P.ws) + (long)ks * M * N;
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const int m = m0 + wm * (BM / WM) + i * MF + Mf::drow(lane, v);
          const int n = n0 + wn * (BN / WN) + j * MF + Mf::dcol(lane);
          if (m < M && n < N) W[(long)m * N + n] = make_float2(accR[i][j][v], accI[i][j][v]);
        }
    return;
  }
  // BF16-plane outputs are staged through LDS so that every lane stores 16
  // contiguous bytes of one plane (the MFMA layout gives a lane 4 rows of one
  // column); column nl of tile row ml sits at stage_col(ml, nl).
  constexpr bool kStaged = 2 * TERMS * BM * BN <= NBUF * STAGE;
  if (TERMS == 2 && BN % 64 == 0 && kStaged && P.out_kind == OUT_PLANES && !P.transC && P.terms == TERMS) {
    // The FP32 sums go through LDS (row ml, position stage_col(ml, nl)); then each thread
    // finishes 8 consecutive outputs of one row with 16-byte reads of the addend planes
    // and of a whole-grid real mask, the epilogue arithmetic of epilogue_value, the BF16
    // split, and one 16-byte store per plane.
    float2* st = reinterpret_cast<float2*>(lds);
    __syncthreads();                               // every wave is done reading operand tiles
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const int ml = wm * (BM / WM) + i * MF + Mf::drow(lane, v);
          const int nl = wn * (BN / WN) + j * MF + Mf::dcol(lane);
          st[ml * BN + stage_col(ml, nl)] = make_float2(accR[i][j][v], accI[i][j][v]);
        }
    __syncthreads();
    const bool empty = nslab == 0;
    const int mk = P.mask_kind;
    const bool vmask = mk == MASK_NONE || ((mk == MASK_REAL_1PI || mk == MASK_REAL_1PI_F32 || mk == MASK_REAL_1PI_U8 ||
                                            mk == MASK_REAL) && P.mm0 == 0 && P.mn0 == 0 && P.mM >= M && P.mN >= N);
    constexpr int CPR = BN / 8;
    GLOBAL unsigned short* C = gout<unsigned short>(P.C);
    for (int e = t; e < BM * CPR; e += NT) {
      const int ml = e / CPR, cc = e % CPR;
      const int m = m0 + ml, n = n0 + cc * 8;      // N is a multiple of 8: chunks are whole
      if (m >= M || n >= N) continue;
      const long off = (long)m * P.ldc + n;
      float av[4][8];                              // addend planes
      if (P.addend && !empty) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const u32x4 w = *reinterpret_cast<const GLOBAL u32x4*>(gin<unsigned short>(P.addend) + q * P.sC + off);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            av[q][2 * k] = bf16_val((unsigned short)(w[k] & 0xFFFFu));
            av[q][2 * k + 1] = bf16_val((unsigned short)(w[k] >> 16));
          }
        }
      }
      double wm8[8];                               // mask values (whole-grid real masks)
      if (vmask && mk != MASK_NONE && !empty) {
        const long mo = (long)m * P.ldm + n;
        if (mk == MASK_REAL_1PI_U8) {
          const u32x2 w = *reinterpret_cast<const GLOBAL u32x2*>(gin<unsigned char>(P.mask) + mo);
#pragma unroll
          for (int k = 0; k < 8; k++) wm8[k] = (double)((w[k >> 2] >> (8 * (k & 3))) & 0xFFu);
        } else if (mk == MASK_REAL_1PI_F32) {
#pragma unroll
          for (int k = 0; k < 8; k++) wm8[k] = (double)gin<float>(P.mask)[mo + k];
        } else {
#pragma unroll
          for (int k = 0; k < 8; k++) wm8[k] = gin<double>(P.mask)[mo + k];
        }
      }
      unsigned short h[4][8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const float2 a = st[ml * BN + stage_col(ml, cc * 8 + k)];
        double vr = 0.0, vi = 0.0;
        if (!empty) {
          if (vmask) {
            vr = P.alpha_re * (double)a.x - P.alpha_im * (double)a.y;
            vi = P.alpha_re * (double)a.y + P.alpha_im * (double)a.x;
            if (P.addend) {
              vr += (double)av[0][k] + (double)av[2][k];
              vi += (double)av[1][k] + (double)av[3][k];
            }
            if (mk == MASK_REAL) {
              vr *= wm8[k]; vi *= wm8[k];
            } else if (mk != MASK_NONE) {          // csim mask: (1+i) w
              const double tr = (vr - vi) * wm8[k];
              vi = (vr + vi) * wm8[k];
              vr = tr;
            }
          } else {
            epilogue_value(P, m, n + k, (double)a.x, (double)a.y, vr, vi);
          }
        }
        float xr = (float)vr, xi = (float)vi;
        h[0][k] = bf16_rne(xr);
        h[1][k] = bf16_rne(xi);
        h[2][k] = bf16_rne(xr - bf16_val(h[0][k]));
        h[3][k] = bf16_rne(xi - bf16_val(h[1][k]));
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
        *reinterpret_cast<GLOBAL u32x4*>(C + q * P.sC + off) =
            u32x4{(unsigned)h[q][0] | ((unsigned)h[q][1] << 16), (unsigned)h[q][2] | ((unsigned)h[q][3] << 16),
                  (unsigned)h[q][4] | ((unsigned)h[q][5] << 16), (unsigned)h[q][6] | ((unsigned)h[q][7] << 16)};
    }
    return;
  }
  constexpr bool kStaged64 = BM * BN * 16 <= 2 * NBUF * STAGE && BN % 64 == 0;
  if (TERMS == 3 && kStaged64 && P.out_kind == OUT_PLANES && !P.transC && P.terms == 3) {
    // Three-term planes, whole-grid real mask or none (the common case): the FP64 sums go
    // through LDS (row ml, position stage_col(ml, nl)), then each thread finishes 8
    // consecutive outputs of one row (epilogue_value's arithmetic with 8-byte mask reads
    // and 16-byte reads of the addend's planes), the 4 threads of a window of 32 columns
    // (4 consecutive lanes: BN / 8 chunks per row, a multiple of 4) agree on its exponent,
    // and each thread stores one 16-byte chunk per plane.  Other epilogues take the
    // general path below.
    const int mk = P.mask_kind;
    const bool vmask = mk == MASK_NONE || ((mk == MASK_REAL_1PI || mk == MASK_REAL_1PI_F32 ||
                                            mk == MASK_REAL_1PI_U8 || mk == MASK_REAL) &&
                                           P.mm0 == 0 && P.mn0 == 0 && P.mM >= M && P.mN >= N);
    if (vmask) {
      f64x2* st = reinterpret_cast<f64x2*>(lds);
      __syncthreads();                             // every wave is done reading operand tiles
#pragma unroll
      for (int i = 0; i < IM; i++)
#pragma unroll
        for (int j = 0; j < JN; j++)
#pragma unroll
          for (int v = 0; v < NV; v++) {
            const int ml = wm * (BM / WM) + i * MF + Mf::drow(lane, v);
            const int nl = wn * (BN / WN) + j * MF + Mf::dcol(lane);
            st[ml * BN + stage_col(ml, nl)] =
                f64x2{(double)accR[i][j][v] + dR[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v],
                      (double)accI[i][j][v] + dI[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v]};
          }
      __syncthreads();
      const bool empty = nslab == 0;
      constexpr int CPR = BN / 8;
      static_assert(!kStaged64 || (CPR % 4 == 0 && NT % 4 == 0 && (BM * CPR) % NT == 0),
                    "a window is 4 consecutive lanes, every lane runs every iteration");
      GLOBAL unsigned short* C = gout<unsigned short>(P.C);
      for (int e = t; e < BM * CPR; e += NT) {     // uniform trip count: BM CPR is a multiple of NT
        const int ml = e / CPR, cc = e % CPR;
        const int m = m0 + ml, n = n0 + cc * 8;    // N is a multiple of 32: windows are whole
        const bool live = m < M && n < N;          // uniform over a window (one row, whole windows)
        const long off = (long)m * P.ldc + n;
        double wm8[8];
        if (live && mk != MASK_NONE && !empty) {
          const long mo = (long)m * P.ldm + n;
          if (mk == MASK_REAL_1PI_U8) {
            const u32x2 w = *reinterpret_cast<const GLOBAL u32x2*>(gin<unsigned char>(P.mask) + mo);
#pragma unroll
            for (int k = 0; k < 8; k++) wm8[k] = (double)((w[k >> 2] >> (8 * (k & 3))) & 0xFFu);
          } else if (mk == MASK_REAL_1PI_F32) {
#pragma unroll
            for (int k = 0; k < 8; k++) wm8[k] = (double)gin<float>(P.mask)[mo + k];
          } else {
#pragma unroll
            for (int k = 0; k < 8; k++) wm8[k] = gin<double>(P.mask)[mo + k];
          }
        }
        // the addend (stored like the output: three-term planes) of the 8 outputs, each
        // value summed hi + mid + lo in FP64 as epilogue_value does
        double ar8[8], ai8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) { ar8[k] = 0.0; ai8[k] = 0.0; }
        if (P.addend && live && !empty) {
          u32x4 aw[6];
#pragma unroll
          for (int q = 0; q < 6; q++)
            aw[q] = *reinterpret_cast<const GLOBAL u32x4*>(gin<unsigned short>(P.addend) + q * P.sC + off);
#pragma unroll
          for (int k = 0; k < 8; k++)
#pragma unroll
            for (int tt = 0; tt < 3; tt++) {
              const unsigned short hr = (unsigned short)(aw[2 * tt][k >> 1] >> (16 * (k & 1)));
              const unsigned short hi = (unsigned short)(aw[2 * tt + 1][k >> 1] >> (16 * (k & 1)));
              ar8[k] += (double)bf16_val(hr);
              ai8[k] += (double)bf16_val(hi);
            }
        }
        double vr[8], vi[8];
        float mx = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const f64x2 a = st[ml * BN + stage_col(ml, cc * 8 + k)];
          double xr = 0.0, xi = 0.0;
          if (live && !empty) {
            xr = P.alpha_re * a.x - P.alpha_im * a.y;
            xi = P.alpha_re * a.y + P.alpha_im * a.x;
            if (P.addend) { xr += ar8[k]; xi += ai8[k]; }
            if (mk == MASK_REAL) {
              xr *= wm8[k]; xi *= wm8[k];
            } else if (mk != MASK_NONE) {          // csim mask: (1+i) w
              const double tr = (xr - xi) * wm8[k];
              xi = (xr + xi) * wm8[k];
              xr = tr;
            }
          }
          vr[k] = xr;
          vi[k] = xi;
          mx = fmaxf(mx, (float)fmax(fabs(xr), fabs(xi)));
        }
        const int ew = window_exponent(window_max<4>(mx));
        unsigned short h[6][8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          unsigned short hr[3], hi[3];
          split3(vr[k], ew, hr);
          split3(vi[k], ew, hi);
#pragma unroll
          for (int tt = 0; tt < 3; tt++) {
            h[2 * tt][k] = hr[tt];
            h[2 * tt + 1][k] = hi[tt];
          }
        }
        if (live) {
#pragma unroll
          for (int q = 0; q < 6; q++)
            *reinterpret_cast<GLOBAL u32x4*>(C + q * P.sC + off) =
                u32x4{(unsigned)h[q][0] | ((unsigned)h[q][1] << 16), (unsigned)h[q][2] | ((unsigned)h[q][3] << 16),
                      (unsigned)h[q][4] | ((unsigned)h[q][5] << 16), (unsigned)h[q][6] | ((unsigned)h[q][7] << 16)};
        }
      }
      return;
    }
  }
  if (TERMS == 3 && P.out_kind == OUT_PLANES) {
    // Three-term planes (the next contraction's operand, K along n): every value of a
    // window of kPlaneWindow consecutive columns of a row gets its leading part on the
    // window's fixed-point grid (planes.hpp).  A row's window lies in one wave: the
    // kPlaneWindow / MF subtiles j of a window, MF lanes each (Mf::dcol).
    constexpr int JW = kPlaneWindow / MF;
    static_assert(TERMS != 3 || (JN % JW == 0 && (BN / WN) % kPlaneWindow == 0), "window within a wave");
    double vr[IM][JN][NV], vi[IM][JN][NV];
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const int m = m0 + wm * (BM / WM) + i * MF + Mf::drow(lane, v);
          const int n = n0 + wn * (BN / WN) + j * MF + Mf::dcol(lane);
          vr[i][j][v] = 0.0; vi[i][j][v] = 0.0;
          if (nslab > 0 && m < M && n < N)
            epilogue_value(P, m, n, dR[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v],
                           dI[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v], vr[i][j][v], vi[i][j][v]);
        }
    int ew[IM][JN / JW][NV];
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int w = 0; w < JN / JW; w++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          float mx = 0.0f;
#pragma unroll
          for (int j = w * JW; j < (w + 1) * JW; j++)
            mx = fmaxf(mx, (float)fmax(fabs(vr[i][j][v]), fabs(vi[i][j][v])));
          ew[i][w][v] = window_exponent(window_max<MF>(mx));
        }
    GLOBAL unsigned short* C = gout<unsigned short>(P.C);
    if (kStaged && !P.transC && P.terms == TERMS) {
      __syncthreads();                             // every wave is done reading operand tiles
#pragma unroll
      for (int i = 0; i < IM; i++)
#pragma unroll
        for (int j = 0; j < JN; j++)
#pragma unroll
          for (int v = 0; v < NV; v++) {
            const int ml = wm * (BM / WM) + i * MF + Mf::drow(lane, v);
            const int nl = wn * (BN / WN) + j * MF + Mf::dcol(lane);
            const int pos = ml * BN + ((((nl >> 3) ^ (((ml >> 2) & 3) << 1))) << 3) + (nl & 7);
            unsigned short hr[3], hi[3];
            split3(vr[i][j][v], ew[i][j / JW][v], hr);
            split3(vi[i][j][v], ew[i][j / JW][v], hi);
#pragma unroll
            for (int tt = 0; tt < 3; tt++) {
              lds[(2 * tt) * BM * BN + pos] = hr[tt];
              lds[(2 * tt + 1) * BM * BN + pos] = hi[tt];
            }
          }
      __syncthreads();
      constexpr int CPR = BN / 8;
      for (int e = t; e < 2 * TERMS * BM * CPR; e += NT) {
        const int q = e / (BM * CPR), rem = e % (BM * CPR), ml = rem / CPR, cc = rem % CPR;
        const int m = m0 + ml, n = n0 + cc * 8;    // N is a multiple of 8: chunks are whole
        if (m < M && n < N)
          *reinterpret_cast<GLOBAL u32x4*>(C + q * P.sC + (long)m * P.ldc + n) =
              *reinterpret_cast<const u32x4*>(lds + q * BM * BN + ml * BN + ((cc ^ (((ml >> 2) & 3) << 1)) << 3));
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const int m = m0 + wm * (BM / WM) + i * MF + Mf::drow(lane, v);
          const int n = n0 + wn * (BN / WN) + j * MF + Mf::dcol(lane);
          if (m < M && n < N)
            store_planes3(C, P.transC ? (long)n * P.ldc + m : (long)m * P.ldc + n, P.sC, vr[i][j][v], vi[i][j][v],
                          ew[i][j / JW][v]);
        }
    return;
  }
  if (TERMS == 2 && kStaged && P.out_kind == OUT_PLANES && !P.transC && P.terms == TERMS) {
    __syncthreads();                               // every wave is done reading operand tiles
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < JN; j++)
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const int ml = wm * (BM / WM) + i * MF + Mf::drow(lane, v);
          const int nl = wn * (BN / WN) + j * MF + Mf::dcol(lane);
          double re = accR[i][j][v], im = accI[i][j][v];
          if (TERMS == 3) { re += dR[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v]; im += dI[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v]; }
          double vr = 0.0, vi = 0.0;
          if (nslab > 0) epilogue_value(P, m0 + ml, n0 + nl, re, im, vr, vi);
          const int pos = ml * BN + ((((nl >> 3) ^ (((ml >> 2) & 3) << 1))) << 3) + (nl & 7);
          float xr = (float)vr, xi = (float)vi;
#pragma unroll
          for (int tt = 0; tt < TERMS; tt++) {
            const unsigned short hr = bf16_rne(xr), hi = bf16_rne(xi);
            lds[(2 * tt) * BM * BN + pos] = hr;
            lds[(2 * tt + 1) * BM * BN + pos] = hi;
            xr -= bf16_val(hr);
            xi -= bf16_val(hi);
          }
        }
    __syncthreads();
    constexpr int CPR = BN / 8;
    GLOBAL unsigned short* C = gout<unsigned short>(P.C);
    for (int e = t; e < 2 * TERMS * BM * CPR; e += NT) {
      const int q = e / (BM * CPR), rem = e % (BM * CPR), ml = rem / CPR, cc = rem % CPR;
      const int m = m0 + ml, n = n0 + cc * 8;      // N is a multiple of 8: chunks are whole
      if (m < M && n < N)
        *reinterpret_cast<GLOBAL u32x4*>(C + q * P.sC + (long)m * P.ldc + n) =
            *reinterpret_cast<const u32x4*>(lds + q * BM * BN + ml * BN + ((cc ^ (((ml >> 2) & 3) << 1)) << 3));
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < IM; i++)
#pragma unroll
    for (int j = 0; j < JN; j++)
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const int m = m0 + wm * (BM / WM) + i * MF + Mf::drow(lane, v);
        const int n = n0 + wn * (BN / WN) + j * MF + Mf::dcol(lane);
        double re = accR[i][j][v], im = accI[i][j][v];
        if (TERMS == 3) { re += dR[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v]; im += dI[TERMS == 3 ? i : 0][TERMS == 3 ? j : 0][v]; }
        epilogue_store(P, m, n, re, im, nslab == 0);
      }
}

// Tile variants of each term count, in the order the tuner tries them.
struct PlanesVariant {
  int bm, bn, threads;
  void (*kernel)(const GemmProb*, int);
  const char* name;    // template arguments: the tuner cache keys its choices on them
};
#define CSX_PLANES(BM, BN, TH, ...) {BM, BN, TH, k_gemm_planes<__VA_ARGS__>, #__VA_ARGS__}
// (the last template argument is the minimum waves per SIMD the registers must allow)
static const PlanesVariant kVariants2[] = {
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, false, 3, 32, 16),
    CSX_PLANES(128, 128, 256, 2, 128, 128, 2, 2, true, 1, 32, 16),
    CSX_PLANES(128, 128, 512, 2, 128, 128, 2, 4, true, 2, 32, 16),
    CSX_PLANES(128, 64, 256, 2, 128, 64, 2, 2, false, 2, 32, 16),
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, true, 2, 32, 16),
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, false, 2, 64, 16),
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, true, 2, 64, 16),
    CSX_PLANES(128, 128, 256, 2, 128, 128, 2, 2, false, 1, 32, 32),
    CSX_PLANES(128, 128, 512, 2, 128, 128, 2, 4, true, 2, 32, 32),
    CSX_PLANES(128, 64, 256, 2, 128, 64, 2, 2, false, 2, 32, 32),
    CSX_PLANES(64, 128, 256, 2, 64, 128, 2, 2, false, 2, 32, 32),
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, true, 2, 32, 16, true),
    CSX_PLANES(128, 128, 512, 2, 128, 128, 2, 4, true, 2, 32, 16, true),
    CSX_PLANES(128, 64, 256, 2, 128, 64, 2, 2, true, 2, 32, 16, true),
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, true, 2, 32, 16, 3),
    CSX_PLANES(64, 64, 256, 2, 64, 64, 2, 2, true, 2, 32, 16, 4),
    CSX_PLANES(128, 64, 256, 2, 128, 64, 2, 2, true, 2, 32, 16, 3),
};
static const PlanesVariant kVariants3[] = {
    CSX_PLANES(64, 64, 256, 3, 64, 64, 2, 2, false, 1, 32, 16),
    CSX_PLANES(128, 64, 512, 3, 128, 64, 4, 2, false, 2, 32, 16),
    CSX_PLANES(64, 64, 256, 3, 64, 64, 2, 2, true, 1, 32, 16),
    CSX_PLANES(64, 64, 256, 3, 64, 64, 2, 2, true, 1, 32, 16, true),
    CSX_PLANES(64, 64, 256, 3, 64, 64, 2, 2, false, 1, 32, 32),
    CSX_PLANES(128, 64, 256, 3, 128, 64, 2, 2, false, 1, 32, 32),
    CSX_PLANES(64, 128, 256, 3, 64, 128, 2, 2, false, 1, 32, 32),
    CSX_PLANES(128, 64, 512, 3, 128, 64, 4, 2, true, 2, 32, 16, true),
    CSX_PLANES(64, 64, 256, 3, 64, 64, 2, 2, true, 1, 32, 16, 3),
    CSX_PLANES(64, 64, 256, 3, 64, 64, 2, 2, true, 1, 32, 32, 3),
};

std::string gemm_variant_names() {
  // the kernel-set revision: split K for every precision (FP64 workspaces for FP64,
  // FP32 and three-term planes)
  std::string s = "splitk-all;f64:";
  for (const auto& v : kVariantsF64) { s += v.name; s += ';'; }
  s += "p2:";
  for (const auto& v : kVariants2) { s += v.name; s += ';'; }
  s += "p3:";
  for (const auto& v : kVariants3) { s += v.name; s += ';'; }
  return s;
}

int planes_variant_count(int terms) {
  return terms == 2 ? (int)(sizeof(kVariants2) / sizeof(kVariants2[0]))
                    : (terms == 3 ? (int)(sizeof(kVariants3) / sizeof(kVariants3[0])) : 0);
}
int planes_variant_tile(int terms, int v, int* bm, int* bn) {
  if (v < 0 || v >= planes_variant_count(terms)) return fail(CSX_ERR_ARG, "planes variant out of range");
  const PlanesVariant& pv = terms == 2 ? kVariants2[v] : kVariants3[v];
  *bm = pv.bm; *bn = pv.bn;
  return CSX_OK;
}
// One thread per output element, n fastest.  Three-term plane outputs take the leading
// part of every value on the fixed-point grid of its window of kPlaneWindow consecutive
// columns (planes.hpp): the windows are whole groups of 32 lanes (N and every problem's
// element offset are multiples of kPlaneWindow, checked at plan time), and every lane of
// a group runs the loop body.
__global__ __launch_bounds__(256) void k_splitk_finish(const GemmProb* __restrict__ probs, int nprob, long total) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e - threadIdx.x % kPlaneWindow < total; e += stride) {
    const bool live = e < total;
    const long el = live ? e : total - 1;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].elem_start <= el) lo = mid; else hi = mid - 1;
    }
    const GemmProb& P = probs[lo];
    const long q = el - P.elem_start, mn = (long)P.M * P.N;
    const int m = (int)(q / P.N), n = (int)(q - (long)m * P.N);
    // a block that is zero and never read (K range lo = -1): its tiles wrote no sums
    // and nothing is stored there
    const bool skip = !live || (P.kr && P.kr[2 * ((long)(m / kKrBlock) * P.kr_ld + n / kKrBlock)] < 0);
    double re = 0, im = 0;
    if (!skip) {
      if (P.ws_f64) {
        const f64x2* W = reinterpret_cast<const f64x2*>(P.ws);
        for (int k = 0; k < P.ksplit; k++) {       // fixed order: deterministic
          const f64x2 w = W[(long)k * mn + q];
          re += w.x;
          im += w.y;
        }
      } else {
        const float2* W = reinterpret_cast<const float2*>(P.ws);
        for (int k = 0; k < P.ksplit; k++) {
          const float2 w = W[(long)k * mn + q];
          re += w.x;
          im += w.y;
        }
      }
    }
    if (P.out_kind == OUT_PLANES && P.terms == 3) {
      // uniform over the whole grid: a launch's problems share out_kind and terms
      double vr = 0.0, vi = 0.0;
      if (!skip) epilogue_value(P, m, n, re, im, vr, vi);
      const int ew = window_exponent(window_max<kPlaneWindow>((float)fmax(fabs(vr), fabs(vi))));
      if (!skip) {
        const long off = P.transC ? (long)n * P.ldc + m : (long)m * P.ldc + n;
        store_planes3(gout<unsigned short>(P.C), off, P.sC, vr, vi, ew);
      }
    } else if (!skip) {
      epilogue_store(P, m, n, re, im);
    }
  }
}
int gemm_splitk_finish(const GemmProb* probs_dev, int n, long total_elems, hipStream_t s) {
  if (total_elems <= 0) return CSX_OK;
  const int blocks = (int)std::min<long>(8192, (total_elems + 255) / 256);
  hipLaunchKernelGGL(k_splitk_finish, dim3(blocks), dim3(256), 0, s, probs_dev, n, total_elems);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

int gemm_bf16planes(const GemmProb* probs_dev, int n, int total_tiles, int terms, int variant, hipStream_t s) {
  if (total_tiles <= 0) return CSX_OK;
  if (variant < 0 || variant >= planes_variant_count(terms)) return fail(CSX_ERR_ARG, "gemm_bf16planes: bad variant");
  const PlanesVariant& pv = terms == 2 ? kVariants2[variant] : kVariants3[variant];
  hipLaunchKernelGGL(pv.kernel, dim3(total_tiles), dim3(pv.threads), 0, s, probs_dev, n);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

// Three-term planes: npix is a multiple of kPlaneWindow (whole windows per column of a
// field whose row count is), and the windows are 32 consecutive lanes of the loop.
__global__ void k_split_planes(const double2* __restrict__ in, unsigned short* __restrict__ out, long npix,
                               long total, int terms) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    long f = i / npix, p = i - f * npix;
    double2 v = in[i];
    int e = 0;
    if (terms == 3) e = window_exponent(window_max<kPlaneWindow>((float)fmax(fabs(v.x), fabs(v.y))));
    store_planes_w(out + f * 2 * terms * npix, p, npix, terms, v.x, v.y, e);
  }
}
int split_c128_to_planes(const void* in, void* out, long npix, int nfield, int terms, hipStream_t s) {
  long total = npix * nfield;
  if (total <= 0) return CSX_OK;
  if (terms == 3 && npix % kPlaneWindow) return fail(CSX_ERR_ARG, "split_c128_to_planes: pixels per field not whole windows");
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
    // |cos|, |sin| <= 1: exponent 0 (unit 2^-8) bounds every window of three-term planes
    if (S.terms > 0) store_planes_w(static_cast<unsigned short*>(S.out), e, total, S.terms, cs, sn, 0);
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
