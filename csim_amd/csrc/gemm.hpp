// Grouped complex GEMM on CDNA4 MFMA: the contraction engine of every matrix
// Fourier transform (csim zoomFft, src/lib/csim_fft.cpp:670-826, restated as
// E_out = s * A_y * E_in * B_x^T).
#pragma once
#include <string>

#include "common.hpp"

namespace csx {

// MASK_COMPLEX_CLIP: complex factor inside the mask rectangle, zero outside it.
// MASK_REAL_1PI_F32 / _U8: MASK_REAL_1PI with the mask stored in float (values exact in
// float) or in bytes (values 0 and 1 only).
enum MaskKind : int {
  MASK_NONE = 0, MASK_COMPLEX = 1, MASK_REAL_1PI = 2, MASK_REAL = 3, MASK_COMPLEX_CLIP = 4, MASK_REAL_1PI_F32 = 5,
  MASK_REAL_1PI_U8 = 6
};
enum OutKind : int { OUT_C128 = 0, OUT_C64 = 1, OUT_PLANES = 2 };

// One problem of a grouped launch:  C(m, n) = alpha * sum_k A(m, k) B(n, k),
// then the epilogue mask.  A and B are K-contiguous (row-major M x K, N x K).
// Element types are complex128 (FP64 kernel) or complex64 (FP32 / split-BF16).
struct GemmProb {
  const void* A;
  const void* B;
  void* C;
  long lda, ldb, ldc;
  int M, N, K;
  int transC;          // 0: C[m*ldc + n]; 1: C[n*ldc + m]
  int out_kind;        // OutKind
  double alpha_re, alpha_im;
  int mask_kind;       // MaskKind; applied at output (m, n)
  const void* mask;    // element (m - mm0, n - mn0) at [(m - mm0) * ldm + (n - mn0)]
  long ldm;
  int mm0, mn0, mM, mN;
  int tiles_m, tiles_n, tile_start;
  // split-BF16 planes: plane q = 2 t + c (term t, c = re/im) of an operand sits
  // q * sX bf16 elements after its base pointer
  long sA, sB, sC;
  int terms;
  // split K: ksplit > 1 blocks per output tile each take a K range (whole slabs) and
  // write their sums to ws ([split][m][n]: complex float for two-term BF16 planes,
  // complex double for FP64, FP32 and three-term planes, ws_f64 = 1); a finishing
  // launch (gemm_splitk_finish) adds the splits in order and applies the epilogue
  int ksplit;
  void* ws;
  int ws_f64;
  long elem_start;     // first output element of this problem among the launch's (finish kernel)
  // optional K range per 64 x 64 block of outputs: int pair [lo, hi) at
  // kr[2 * (mb * kr_ld + nb)] bounds the nonzero products of every output of block
  // (mb, nb); a tile runs the whole K slabs of the union over its blocks (the
  // products it skips are exact zeros).  lo >= hi: the block's outputs are zero;
  // lo = -1 (hi = 0): zero and never read, a tile of such blocks writes nothing.
  const int* kr;
  int kr_ld;
  unsigned kr_tag;     // digest of the table (tile tuning keys on it)
  // optional addend stored like C (out_kind, ldc, sC, terms): C = mask (alpha sum + addend)
  const void* addend;
};
constexpr int kKrBlock = 64;

// Tile of the FP64 / FP32 MFMA kernels (64 x 64 outputs, 4 waves of 32 x 32).
constexpr int kTileM = 64, kTileN = 64;

// Fills tiles_m/tiles_n/tile_start (and elem_start); returns the total tile count
// (ksplit tiles per output tile when ksplit > 1).
int plan_tiles(GemmProb* probs, int n, int tile_m, int tile_n);
// Adds the K splits of every problem of a split-K launch and applies the epilogue;
// total_elems = sum of M N over the problems.
int gemm_splitk_finish(const GemmProb* probs_dev, int n, long total_elems, hipStream_t s);

// probs_dev: device copy of the problem array (n entries), total_tiles from plan_tiles.
int gemm_f64(const GemmProb* probs_dev, int n, int total_tiles, hipStream_t s, int variant = 0);
// FP64 tile variants (variant 0: the 64 x 64 single-buffered kernel)
int f64_variant_count();
int f64_variant_tile(int variant, int* bm, int* bn);
int gemm_f32(const GemmProb* probs_dev, int n, int total_tiles, hipStream_t s);
// Operands already split into BF16 planes (A, B K-contiguous per plane, K, lda/ldb
// and plane strides multiples of 8); output complex128/complex64 or planes
// (OUT_PLANES).  `variant` selects a tile shape (planes_variant_tile gives the
// tile the problem list must be planned with).
int planes_variant_count(int terms);
// Every tile variant of every precision, in table order (the tuner cache checks it).
std::string gemm_variant_names();
int planes_variant_tile(int terms, int variant, int* bm, int* bn);
int gemm_bf16planes(const GemmProb* probs_dev, int n, int total_tiles, int terms, int variant, hipStream_t s);
// complex128 -> `terms` BF16 planes per complex element (hi, mid, lo of re and im),
// field f's planes at out + f * 2 * terms * npix, plane q at + q * npix.
int split_c128_to_planes(const void* in, void* out, long npix, int nfield, int terms, hipStream_t s);

// DFT kernel matrix M[o][i] = exp(-2 pi i (xo0 + o dxo)(xi0 + i dxi) / lf), one
// per (problem) entry, FP64 phase argument reduced to [-1/2, 1/2] cycles.
struct DftSpec {
  void* out;          // complex128 or complex64 (nout x nin, row-major)
  int nout, nin;
  double xo0, dxo, xi0, dxi, lf;
  int out_c64;
  int terms;          // > 0: write `terms` BF16 planes (plane stride nout * nin) instead
};
int dft_matrices(const DftSpec* specs_dev, int n, int max_elems, hipStream_t s);

// complex128 <-> complex64 conversions.
int convert_c128_to_c64(const void* in, void* out, size_t n, hipStream_t s);
int convert_c64_to_c128(const void* in, void* out, size_t n, hipStream_t s);

}  // namespace csx
