// Linearised hex-response objective (optHexRespObjective::execute,
// src/optimization/optObjective.cpp:24-69) as a batched complex GEMM on MFMA:
// per wavelength slice s, E_s = M_s * [exp(i a_s X); 1] (the reference's
// zgemm, :46), pixEnergy = sum_s |E_s|^2, objective = mean (or max) / calib.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "phase.hpp"
#include "gemm.hpp"
#include "mft.hpp"

namespace csx {

__device__ __forceinline__ void put(double2* p, double x, double y) { *p = make_double2(x, y); }
__device__ __forceinline__ void put(float2* p, double x, double y) { *p = make_float2((float)x, (float)y); }

// Mt[s][p][i] = M[s](p, i)   (M column-major npix x ncol; Mt K-contiguous with row
// stride ld >= ncol, the columns i in [ncol, ld) zero)
template <typename C2>
__global__ void k_resp_transpose(const double2* __restrict__ M, C2* __restrict__ Mt, int npix, int ncol, int ld,
                                 int nsl) {
  __shared__ double2 tile[32][33];
  int s = blockIdx.z;
  const double2* Ms = M + (size_t)s * npix * ncol;
  C2* Ts = Mt + (size_t)s * npix * ld;
  int p0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int i = i0 + k, p = p0 + threadIdx.x;
    tile[k][threadIdx.x] = (i < ncol && p < npix) ? Ms[(size_t)i * npix + p] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int p = p0 + k, i = i0 + threadIdx.x;
    if (i < ld && p < npix) {
      double2 v = tile[threadIdx.x][k];
      put(&Ts[(size_t)p * ld + i], v.x, v.y);
    }
  }
}

// R[s][j][i] = exp(i a_s X[j][i]) (i < nopt), 1 for the appended column i = nopt, 0 for
// the padding columns i in (nopt, ld).
template <typename C2>
__global__ void k_resp_phases(const double* __restrict__ X, int nopt, int ld, int ncand,
                              const double* __restrict__ a_s, int nsl, C2* __restrict__ R) {
#pragma clang fp contract(off)
  long total = (long)nsl * ncand * ld;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    int i = (int)(e % ld);
    long sj = e / ld;
    int j = (int)(sj % ncand), s = (int)(sj / ncand);
    if (i > nopt) {
      put(&R[e], 0.0, 0.0);
      continue;
    }
    double x = i < nopt ? X[(long)j * nopt + i] : 0.0;
    double ph = a_s[s] * x;
    double sn, cs;
    sincos_phase(ph, &sn, &cs);
    put(&R[e], cs, sn);
  }
}

// obj[j] = reduce_p sum_s |E_s(j, p)|^2 / calib; one block per candidate.
__global__ void k_resp_reduce(const double2* __restrict__ E, int ncand, int npix, int nsl, int use_max,
                              double calib, double* __restrict__ obj) {
  __shared__ double red[32];
  int j = blockIdx.x;
  double acc = use_max ? -INFINITY : 0.0;
  for (int p = threadIdx.x; p < npix; p += blockDim.x) {
    double e = 0;
    for (int s = 0; s < nsl; s++) {
      double2 v = E[((size_t)s * ncand + j) * npix + p];
      e += v.x * v.x + v.y * v.y;
    }
    acc = use_max ? fmax(acc, e) : acc + e;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(acc, o, 64);
    acc = use_max ? fmax(acc, w) : acc + w;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = use_max ? fmax(r, red[w]) : r + red[w];
    obj[j] = (use_max ? r : r / npix) / calib;
  }
}

}  // namespace csx

using namespace csx;

extern "C" int csx_hex_resp_objective(csx_ctx* ctx, const void* d_M, int npix, int ncol, int nslices,
                                      const double* lambdas, double phase_sign, double calib, const double* d_X,
                                      int nopt, int ncand, int objective, double* d_obj, csx_precision p) {
  if (!ctx || !d_M || !lambdas || !d_X || !d_obj) return fail(CSX_ERR_ARG, "csx_hex_resp_objective: null argument");
  if (npix <= 0 || ncol != nopt + 1 || nslices <= 0 || ncand <= 0 || calib == 0)
    return fail(CSX_ERR_ARG, "csx_hex_resp_objective: bad sizes (ncol must be nopt + 1)");
  CSX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // split-BF16 precisions: the operands go through complex128 (K padded with zeros to
  // whole windows of 32) and are split into BF16 planes, K-contiguous per plane
  const int T = terms_of(p);
  const bool c64 = p == CSX_FP32;
  const int ld = T ? (ncol + 31) / 32 * 32 : ncol;
  const size_t esz = c64 ? sizeof(float2) : sizeof(double2);
  // workspace: Mt, R (esz), E (complex128), a_s, then the BF16 planes of Mt and R
  size_t nMt = (size_t)nslices * npix * ld, nR = (size_t)nslices * ncand * ld, nE = (size_t)nslices * ncand * npix;
  size_t off_R = Arena::round(nMt * esz), off_E = off_R + Arena::round(nR * esz);
  size_t off_a = off_E + Arena::round(nE * sizeof(double2));
  size_t off_pM = off_a + Arena::round(sizeof(double) * nslices);
  size_t off_pR = off_pM + Arena::round(T ? nMt * 4 * T : 0);
  size_t total = off_pR + Arena::round(T ? nR * 4 * T : 0);
  CSX_TRY(ctx->work.reserve(total));
  char* w = static_cast<char*>(ctx->work.p);
  std::vector<double> a(nslices);
  for (int i = 0; i < nslices; i++) a[i] = phase_sign * 4 * M_PI / lambdas[i];   // optObjective.cpp:38
  CSX_HIP(hipMemcpyAsync(w + off_a, a.data(), sizeof(double) * nslices, hipMemcpyHostToDevice, s));
  dim3 tb(32, 8), tg((npix + 31) / 32, (ld + 31) / 32, nslices);
  int blocks = (int)std::min<size_t>(4096, (nR + 255) / 256);
  if (c64) {
    hipLaunchKernelGGL(k_resp_transpose<float2>, tg, tb, 0, s, (const double2*)d_M, (float2*)w, npix, ncol, ld,
                       nslices);
    hipLaunchKernelGGL(k_resp_phases<float2>, dim3(blocks), dim3(256), 0, s, d_X, nopt, ld, ncand,
                       (const double*)(w + off_a), nslices, (float2*)(w + off_R));
  } else {
    hipLaunchKernelGGL(k_resp_transpose<double2>, tg, tb, 0, s, (const double2*)d_M, (double2*)w, npix, ncol, ld,
                       nslices);
    hipLaunchKernelGGL(k_resp_phases<double2>, dim3(blocks), dim3(256), 0, s, d_X, nopt, ld, ncand,
                       (const double*)(w + off_a), nslices, (double2*)(w + off_R));
  }
  CSX_HIP(hipGetLastError());
  if (T) {
    // one "field" per slice: its rows of ld values hold whole windows of 32
    CSX_TRY(split_c128_to_planes(w, w + off_pM, (long)npix * ld, nslices, T, s));
    CSX_TRY(split_c128_to_planes(w + off_R, w + off_pR, (long)ncand * ld, nslices, T, s));
  }
  // E_s(j, p) = sum_i R_s[j][i] Mt_s[p][i]
  std::vector<GemmProb> pr(nslices);
  for (int sl = 0; sl < nslices; sl++) {
    GemmProb& g = pr[sl];
    std::memset(&g, 0, sizeof(g));
    if (T) {
      const size_t pa = (size_t)ncand * ld, pb = (size_t)npix * ld;   // plane strides (BF16 elements)
      g.A = reinterpret_cast<unsigned short*>(w + off_pR) + (size_t)sl * 2 * T * pa;
      g.B = reinterpret_cast<unsigned short*>(w + off_pM) + (size_t)sl * 2 * T * pb;
      g.sA = (long)pa; g.sB = (long)pb;
      g.terms = T;
    } else {
      g.A = w + off_R + (size_t)sl * ncand * ld * esz;
      g.B = w + (size_t)sl * npix * ld * esz;
    }
    g.lda = ld; g.ldb = ld;
    g.C = w + off_E + (size_t)sl * ncand * npix * sizeof(double2); g.ldc = npix;
    g.M = ncand; g.N = npix; g.K = ld;
    g.transC = 0; g.out_kind = OUT_C128;
    g.alpha_re = 1.0; g.alpha_im = 0.0;
    g.mask_kind = MASK_NONE;
  }
  int tm = kTileM, tn = kTileN;
  if (T) CSX_TRY(planes_variant_tile(T, 0, &tm, &tn));
  int tiles = plan_tiles(pr.data(), nslices, tm, tn);
  CSX_TRY(ctx->small.reserve(sizeof(GemmProb) * nslices));
  CSX_HIP(hipMemcpyAsync(ctx->small.p, pr.data(), sizeof(GemmProb) * nslices, hipMemcpyHostToDevice, s));
  CSX_TRY(launch_gemm(ctx->small.as<GemmProb>(), nslices, tiles, p, s));
  hipLaunchKernelGGL(k_resp_reduce, dim3(ncand), dim3(256), 0, s, (const double2*)(w + off_E), ncand, npix, nslices,
                     objective == 1 ? 1 : 0, calib, d_obj);
  CSX_HIP(hipGetLastError());
  CSX_HIP(hipStreamSynchronize(s));
  return CSX_OK;
}
