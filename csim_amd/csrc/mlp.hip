// MLP surrogate (no reference counterpart): n_in -> h1 -> h2 -> n_out with GELU,
// trained by the library's own HIP kernels.  Master weights FP32; every product
// rounds its operands to BF16 (round-to-nearest-even) and accumulates in FP32
// on v_mfma_f32_16x16x32_bf16 (mode kBF16), or stays FP32 on
// v_mfma_f32_16x16x4f32 (mode kF32, used to check gradients tightly).  Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"

namespace csx {

enum { kBF16 = 0, kF32 = 1 };
enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2 };

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MlpGemm {
  const float* A; int lda; int transA;   // op(A) is M x K
  const float* B; int ldb; int transB;   // op(B) is K x N
  float* C; int ldc;                     // M x N
  float* C2;                             // GELU output (EPI_BIAS_GELU)
  const float* bias;                     // N
  int M, N, K, epi;
  // split K (EPI_STORE only): blockIdx.z sums K range [z kchunk, (z + 1) kchunk) into
  // ws[z][M][N]; k_splitk_sum adds the ranges in order into C
  int kchunk;
  float* ws;
};

__device__ __forceinline__ float opA(const MlpGemm& g, int m, int k, int kend) {
  if (m >= g.M || k >= kend) return 0.f;
  return g.transA ? g.A[(long)k * g.lda + m] : g.A[(long)m * g.lda + k];
}
__device__ __forceinline__ float opB(const MlpGemm& g, int k, int n, int kend) {
  if (n >= g.N || k >= kend) return 0.f;
  return g.transB ? g.B[(long)n * g.ldb + k] : g.B[(long)k * g.ldb + n];
}
__device__ __forceinline__ float gelu_f(float u) { return 0.5f * u * (1.0f + erff(u * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_df(float u) {
  return 0.5f * (1.0f + erff(u * 0.70710678118654752f)) + u * expf(-0.5f * u * u) * 0.39894228040143268f;
}

// 64x64 tile, 4 waves of 32x32; K staged 32 at a time in LDS (fp32, rounded to
// bf16 on the fragment read in kBF16 mode).
template <int MODE>
__global__ __launch_bounds__(256) void k_mlp_gemm(MlpGemm g) {
  __shared__ float As[64][33];
  __shared__ float Bs[64][33];   // Bs[n][k]
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wm = wave >> 1, wn = wave & 1;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0, 0, 0, 0};
  // Each thread moves 8 values of each operand per 32-deep K step: the index that is
  // contiguous in memory runs fastest over the threads, so each wave's loads are
  // coalesced whichever way the operand is stored.  The next step's values are loaded
  // into registers while the current step's MFMAs run.
  const int kbeg = g.ws ? blockIdx.z * g.kchunk : 0, kend = g.ws ? min(g.K, kbeg + g.kchunk) : g.K;
  float av[8], bv[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int e = t + i * 256;
      const int ra = g.transA ? (e & 63) : (e >> 5), ka = g.transA ? (e >> 6) : (e & 31);
      const int rb = g.transB ? (e >> 5) : (e & 63), kb = g.transB ? (e & 31) : (e >> 6);
      av[i] = opA(g, m0 + ra, k0 + ka, kend);
      bv[i] = opB(g, k0 + kb, n0 + rb, kend);
    }
  };
  gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int e = t + i * 256;
      const int ra = g.transA ? (e & 63) : (e >> 5), ka = g.transA ? (e >> 6) : (e & 31);
      const int rb = g.transB ? (e >> 5) : (e & 63), kb = g.transB ? (e & 31) : (e >> 6);
      As[ra][ka] = av[i];
      Bs[rb][kb] = bv[i];
    }
    __syncthreads();
    if (k0 + 32 < kend) gload(k0 + 32);
    if (MODE == kBF16) {
      const int fr = lane & 15, fk = 8 * (lane >> 4);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
          a[i][e] = (__bf16)As[wm * 32 + i * 16 + fr][fk + e];
          b[i][e] = (__bf16)Bs[wn * 32 + i * 16 + fr][fk + e];
        }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    } else {
#pragma unroll
      for (int kk = 0; kk < 32; kk += 4) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(As[wm * 32 + i * 16 + (lane & 15)][kk + (lane >> 4)],
                                                             Bs[wn * 32 + j * 16 + (lane & 15)][kk + (lane >> 4)],
                                                             acc[i][j], 0, 0, 0);
      }
    }
  }
  // accumulator layout of both MFMA forms: D[4(l/16) + v][l%16]
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
        int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        if (g.ws) {
          g.ws[((long)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][v];
          continue;
        }
        float x = acc[i][j][v];
        if (g.epi != EPI_STORE) x += g.bias[n];
        g.C[(long)m * g.ldc + n] = x;
        if (g.epi == EPI_BIAS_GELU) g.C2[(long)m * g.ldc + n] = gelu_f(x);
      }
}

// C[m][n] = sum over the K ranges z of ws[z][m][n], in order
__global__ void k_splitk_sum(const float* __restrict__ ws, int nsplit, int M, int N, float* __restrict__ C, int ldc) {
  const long total = (long)M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    float x = 0.f;
    for (int z = 0; z < nsplit; z++) x += ws[z * total + e];
    C[(e / N) * ldc + e % N] = x;
  }
}

static int launch(const MlpGemm& g, int mode, hipStream_t s) {
  const int nsplit = g.ws ? (g.K + g.kchunk - 1) / g.kchunk : 1;
  dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, nsplit);
  if (mode == kBF16) hipLaunchKernelGGL(k_mlp_gemm<kBF16>, grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL(k_mlp_gemm<kF32>, grid, dim3(256), 0, s, g);
  CSX_HIP(hipGetLastError());
  if (g.ws) {
    const long total = (long)g.M * g.N;
    hipLaunchKernelGGL(k_splitk_sum, dim3((int)std::min<long>(1024, (total + 255) / 256)), dim3(256), 0, s, g.ws,
                       nsplit, g.M, g.N, g.C, g.ldc);
    CSX_HIP(hipGetLastError());
  }
  return CSX_OK;
}

// dY = 2 (y - t) / (n * nout); per-row loss terms summed by the host
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, long total, float scale,
                      float* __restrict__ dy, double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    float d = y[i] - t[i];
    dy[i] = 2.0f * d * scale;
    acc += (double)d * d * scale;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
// dH = dA * gelu'(H)
__global__ void k_gelu_bwd(float* __restrict__ d, const float* __restrict__ h, long total) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    d[i] *= gelu_df(h[i]);
}
// db[j] = sum_i d[i][j]: a block per 64 columns, 16 row groups of 64 threads, group q
// summing rows q, q + 16, ... in FP64 (loads 16 deep in flight); the groups' sums are
// then added in a fixed order, so the result does not depend on scheduling
constexpr int kColsumGroups = 16;
__global__ __launch_bounds__(64 * kColsumGroups) void k_colsum(const float* __restrict__ d, int n, int m,
                                                               float* __restrict__ out) {
  __shared__ double part[kColsumGroups][64];
  const int c = threadIdx.x & 63, q = threadIdx.x >> 6, j = blockIdx.x * 64 + c;
  double acc = 0;
  if (j < m) {
#pragma unroll 16
    for (int i = q; i < n; i += kColsumGroups) acc += d[(long)i * m + j];
  }
  part[q][c] = acc;
  __syncthreads();
  if (q == 0 && j < m) {
    double s = 0;
    for (int g = 0; g < kColsumGroups; g++) s += part[g][c];
    out[j] = (float)s;
  }
}
// Adam (Kingma & Ba 2015) on fp32 master weights
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float gi = g[i];
    float mi = b1 * m[i] + (1 - b1) * gi;
    float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Surrogate input embedding: row (c, l) = [cos phi_h, sin phi_h (h < nhex), (lambda_l - lc) / ls],
// phi_h = 2 pi (n_index - 1) t[c][h] / lambda_l in FP64.
__global__ void k_phasor_features(const double* __restrict__ t, int ncand, int nhex, const double* __restrict__ lam,
                                  int nlam, double n_index, double lc, double ls, float* __restrict__ out) {
  const int row = blockIdx.x;            // c * nlam + l
  const int c = row / nlam, l = row - c * nlam;
  const double k = 2.0 * M_PI * (n_index - 1.0) / lam[l];
  const int w = 2 * nhex + 1;
  float* o = out + (long)row * w;
  for (int h = threadIdx.x; h < nhex; h += blockDim.x) {
    double sn, cs;
    sincos(k * t[(long)c * nhex + h], &sn, &cs);
    o[h] = (float)cs;
    o[nhex + h] = (float)sn;
  }
  if (threadIdx.x == 0) o[2 * nhex] = (float)((lam[l] - lc) / ls);
}

}  // namespace csx

using namespace csx;

struct csx_mlp {
  csx_ctx* ctx = nullptr;
  int nin = 0, nh1 = 0, nh2 = 0, nout = 0;
  int mode = kBF16;
  long nparam = 0;
  long off[7] = {0};            // W1 b1 W2 b2 W3 b3 end (float offsets)
  DevBuf P, G, Mom, Vel;        // params, grads, Adam moments
  DevBuf act;                   // activations for the largest batch seen
  int act_batch = 0;
  DevBuf partial;
  DevBuf ws;                    // split-K partial sums of the weight gradients
  long step = 0;
  float* p(int i) { return P.as<float>() + off[i]; }
  float* g(int i) { return G.as<float>() + off[i]; }
};

static int mlp_reserve(csx_mlp* m, int n) {
  if (n <= m->act_batch) return CSX_OK;
  size_t per = (size_t)2 * m->nh1 + 2 * m->nh2 + 2 * m->nout + m->nh1 + m->nh2;
  CSX_TRY(m->act.reserve(sizeof(float) * per * n));
  m->act_batch = n;
  return CSX_OK;
}

// forward into the activation buffers: h1 a1 h2 a2 y
static int mlp_fwd(csx_mlp* m, const float* x, int n, float** outs) {
  CSX_TRY(mlp_reserve(m, n));
  hipStream_t s = m->ctx->stream;
  float* h1 = m->act.as<float>();
  float* a1 = h1 + (size_t)n * m->nh1;
  float* h2 = a1 + (size_t)n * m->nh1;
  float* a2 = h2 + (size_t)n * m->nh2;
  float* y = a2 + (size_t)n * m->nh2;
  MlpGemm g{};
  g = MlpGemm{x, m->nin, 0, m->p(0), m->nh1, 0, h1, m->nh1, a1, m->p(1), n, m->nh1, m->nin, EPI_BIAS_GELU};
  CSX_TRY(launch(g, m->mode, s));
  g = MlpGemm{a1, m->nh1, 0, m->p(2), m->nh2, 0, h2, m->nh2, a2, m->p(3), n, m->nh2, m->nh1, EPI_BIAS_GELU};
  CSX_TRY(launch(g, m->mode, s));
  g = MlpGemm{a2, m->nh2, 0, m->p(4), m->nout, 0, y, m->nout, nullptr, m->p(5), n, m->nout, m->nh2, EPI_BIAS};
  CSX_TRY(launch(g, m->mode, s));
  outs[0] = h1; outs[1] = a1; outs[2] = h2; outs[3] = a2; outs[4] = y;
  return CSX_OK;
}

// Weight gradients contract over the batch (K = n) into few output tiles: a long K
// range is split so that the launch has enough blocks (at least 128-deep ranges).
static int launch_wgrad(csx_mlp* m, MlpGemm g, hipStream_t s) {
  const int tiles = ((g.M + 63) / 64) * ((g.N + 63) / 64);
  int nsplit = std::min(8, std::max(1, g.K / 128));
  while (nsplit > 1 && tiles * (nsplit / 2) >= 256) nsplit /= 2;
  if (nsplit > 1) {
    g.kchunk = ((g.K + nsplit - 1) / nsplit + 31) / 32 * 32;
    CSX_TRY(m->ws.reserve(sizeof(float) * (size_t)((g.K + g.kchunk - 1) / g.kchunk) * g.M * g.N));
    g.ws = m->ws.as<float>();
  }
  return launch(g, m->mode, s);
}

// loss + gradients into m->G
static int mlp_bwd(csx_mlp* m, const float* x, const float* t, int n, double* loss) {
  float* o[5];
  CSX_TRY(mlp_fwd(m, x, n, o));
  hipStream_t s = m->ctx->stream;
  float* dy = o[4] + (size_t)n * m->nout;
  float* d2 = dy + (size_t)n * m->nout;
  float* d1 = d2 + (size_t)n * m->nh2;
  long total = (long)n * m->nout;
  int blocks = (int)std::min<long>(1024, (total + 255) / 256);
  CSX_TRY(m->partial.reserve(sizeof(double) * blocks));
  hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, s, o[4], t, total, 1.0f / (float)total, dy,
                     m->partial.as<double>());
  CSX_HIP(hipGetLastError());
  // layer 3: dW3 = a2^T dy, db3 = colsum dy, d2 = dy W3^T * gelu'(h2)
  MlpGemm g{};
  g = MlpGemm{o[3], m->nh2, 1, dy, m->nout, 0, m->g(4), m->nout, nullptr, nullptr, m->nh2, m->nout, n, EPI_STORE};
  CSX_TRY(launch_wgrad(m, g, s));
  hipLaunchKernelGGL(k_colsum, dim3((m->nout + 63) / 64), dim3(64 * kColsumGroups), 0, s, dy, n, m->nout, m->g(5));
  g = MlpGemm{dy, m->nout, 0, m->p(4), m->nout, 1, d2, m->nh2, nullptr, nullptr, n, m->nh2, m->nout, EPI_STORE};
  CSX_TRY(launch(g, m->mode, s));
  hipLaunchKernelGGL(k_gelu_bwd, dim3((int)std::min<long>(4096, ((long)n * m->nh2 + 255) / 256)), dim3(256), 0, s, d2,
                     o[2], (long)n * m->nh2);
  // layer 2
  g = MlpGemm{o[1], m->nh1, 1, d2, m->nh2, 0, m->g(2), m->nh2, nullptr, nullptr, m->nh1, m->nh2, n, EPI_STORE};
  CSX_TRY(launch_wgrad(m, g, s));
  hipLaunchKernelGGL(k_colsum, dim3((m->nh2 + 63) / 64), dim3(64 * kColsumGroups), 0, s, d2, n, m->nh2, m->g(3));
  g = MlpGemm{d2, m->nh2, 0, m->p(2), m->nh2, 1, d1, m->nh1, nullptr, nullptr, n, m->nh1, m->nh2, EPI_STORE};
  CSX_TRY(launch(g, m->mode, s));
  hipLaunchKernelGGL(k_gelu_bwd, dim3((int)std::min<long>(4096, ((long)n * m->nh1 + 255) / 256)), dim3(256), 0, s, d1,
                     o[0], (long)n * m->nh1);
  // layer 1
  g = MlpGemm{x, m->nin, 1, d1, m->nh1, 0, m->g(0), m->nh1, nullptr, nullptr, m->nin, m->nh1, n, EPI_STORE};
  CSX_TRY(launch_wgrad(m, g, s));
  hipLaunchKernelGGL(k_colsum, dim3((m->nh1 + 63) / 64), dim3(64 * kColsumGroups), 0, s, d1, n, m->nh1, m->g(1));
  CSX_HIP(hipGetLastError());
  if (!loss) return CSX_OK;   // no read-back: the step stays asynchronous
  std::vector<double> part(blocks);
  CSX_HIP(hipMemcpyAsync(part.data(), m->partial.p, sizeof(double) * blocks, hipMemcpyDeviceToHost, s));
  CSX_HIP(hipStreamSynchronize(s));
  double L = 0;
  for (double v : part) L += v;
  *loss = L;
  return CSX_OK;
}

extern "C" {

int csx_hex_phasor_features(csx_ctx* ctx, const double* d_thickness, int ncand, int nhex, const double* lambdas,
                            int nlam, double n_index, double lam_center, double lam_scale, float* d_out) {
  if (!ctx || !d_thickness || !lambdas || !d_out || ncand <= 0 || nhex <= 0 || nlam <= 0 || lam_scale == 0)
    return fail(CSX_ERR_ARG, "csx_hex_phasor_features: bad arguments");
  CSX_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  CSX_TRY(ctx->small.reserve(sizeof(double) * nlam));
  CSX_HIP(hipMemcpyAsync(ctx->small.p, lambdas, sizeof(double) * nlam, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_phasor_features, dim3(ncand * nlam), dim3(256), 0, s, d_thickness, ncand, nhex,
                     ctx->small.as<double>(), nlam, n_index, lam_center, lam_scale, d_out);
  CSX_HIP(hipGetLastError());
  CSX_HIP(hipStreamSynchronize(s));
  return CSX_OK;
}

int csx_mlp_create(csx_ctx* ctx, int n_in, int n_h1, int n_h2, int n_out, uint64_t seed, csx_mlp** out) {
  if (!ctx || !out || n_in <= 0 || n_h1 <= 0 || n_h2 <= 0 || n_out <= 0) return fail(CSX_ERR_ARG, "csx_mlp_create: bad arguments");
  *out = nullptr;
  CSX_HIP(hipSetDevice(ctx->device));
  csx_mlp* m = new (std::nothrow) csx_mlp();
  if (!m) return fail(CSX_ERR_NOMEM, "mlp allocation failed");
  m->ctx = ctx;
  m->nin = n_in; m->nh1 = n_h1; m->nh2 = n_h2; m->nout = n_out;
  if (const char* e = std::getenv("CSX_MLP_FP32")) m->mode = (e[0] == '1') ? kF32 : kBF16;
  long sizes[6] = {(long)n_in * n_h1, n_h1, (long)n_h1 * n_h2, n_h2, (long)n_h2 * n_out, n_out};
  for (int i = 0; i < 6; i++) m->off[i + 1] = m->off[i] + sizes[i];
  m->nparam = m->off[6];
  int rc = CSX_OK;
  if ((rc = m->P.reserve(sizeof(float) * m->nparam)) || (rc = m->G.reserve(sizeof(float) * m->nparam)) ||
      (rc = m->Mom.reserve(sizeof(float) * m->nparam)) || (rc = m->Vel.reserve(sizeof(float) * m->nparam))) {
    delete m;
    return rc;
  }
  // He-style init from Philox normals, zero biases
  std::vector<double> nrm(m->nparam);
  csx_philox_fill(seed, 0x4d4c50, nrm.data(), nrm.size(), 1);
  std::vector<float> h(m->nparam, 0.f);
  int fan[3] = {n_in, n_h1, n_h2};
  for (int l = 0; l < 3; l++) {
    float sc = std::sqrt(2.0f / fan[l]);
    for (long i = m->off[2 * l]; i < m->off[2 * l + 1]; i++) h[i] = (float)nrm[i] * sc;
  }
  hipError_t e1 = hipMemcpy(m->P.p, h.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice);
  hipError_t e2 = hipMemset(m->Mom.p, 0, sizeof(float) * m->nparam);
  hipError_t e3 = hipMemset(m->Vel.p, 0, sizeof(float) * m->nparam);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) { delete m; return fail(CSX_ERR_HIP, "mlp init copy failed"); }
  *out = m;
  return CSX_OK;
}

void csx_mlp_destroy(csx_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
}

size_t csx_mlp_param_count(csx_mlp* m) { return m ? (size_t)m->nparam : 0; }

int csx_mlp_get_params(csx_mlp* m, double* flat) {
  if (!m || !flat) return fail(CSX_ERR_ARG, "null argument");
  std::vector<float> h(m->nparam);
  CSX_HIP(hipMemcpy(h.data(), m->P.p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  for (long i = 0; i < m->nparam; i++) flat[i] = h[i];
  return CSX_OK;
}

int csx_mlp_set_params(csx_mlp* m, const double* flat) {
  if (!m || !flat) return fail(CSX_ERR_ARG, "null argument");
  std::vector<float> h(m->nparam);
  for (long i = 0; i < m->nparam; i++) h[i] = (float)flat[i];
  CSX_HIP(hipMemcpy(m->P.p, h.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  // new parameters start a new optimisation: Adam moments and step count restart
  std::fill(h.begin(), h.end(), 0.0f);
  CSX_HIP(hipMemcpy(m->Mom.p, h.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  CSX_HIP(hipMemcpy(m->Vel.p, h.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  m->step = 0;
  return CSX_OK;
}

int csx_mlp_forward(csx_mlp* m, const float* d_x, int n, float* d_y) {
  if (!m || !d_x || !d_y || n <= 0) return fail(CSX_ERR_ARG, "csx_mlp_forward: bad arguments");
  CSX_HIP(hipSetDevice(m->ctx->device));
  float* o[5];
  CSX_TRY(mlp_fwd(m, d_x, n, o));
  CSX_HIP(hipMemcpyAsync(d_y, o[4], sizeof(float) * (size_t)n * m->nout, hipMemcpyDeviceToDevice, m->ctx->stream));
  CSX_HIP(hipStreamSynchronize(m->ctx->stream));
  return CSX_OK;
}

int csx_mlp_loss_grad(csx_mlp* m, const float* d_x, const float* d_t, int n, double* loss, double* grad_flat) {
  if (!m || !d_x || !d_t || n <= 0) return fail(CSX_ERR_ARG, "csx_mlp_loss_grad: bad arguments");
  CSX_HIP(hipSetDevice(m->ctx->device));
  CSX_TRY(mlp_bwd(m, d_x, d_t, n, loss));
  if (grad_flat) {
    std::vector<float> h(m->nparam);
    CSX_HIP(hipMemcpy(h.data(), m->G.p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
    for (long i = 0; i < m->nparam; i++) grad_flat[i] = h[i];
  }
  return CSX_OK;
}

int csx_mlp_train_step(csx_mlp* m, const float* d_x, const float* d_t, int n, double lr, double* loss) {
  if (!m || !d_x || !d_t || n <= 0) return fail(CSX_ERR_ARG, "csx_mlp_train_step: bad arguments");
  CSX_HIP(hipSetDevice(m->ctx->device));
  CSX_TRY(mlp_bwd(m, d_x, d_t, n, loss));
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  int blocks = (int)std::min<long>(4096, (m->nparam + 255) / 256);
  hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, m->ctx->stream, m->P.as<float>(), m->G.as<float>(),
                     m->Mom.as<float>(), m->Vel.as<float>(), m->nparam, (float)lr, b1, b2, 1e-8f, c1, c2);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

}  // extern "C"
