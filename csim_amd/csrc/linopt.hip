// Normal equations of csim's linearised field-conjugation step
// (linearOptimizer::compute_jacobian / compute_objective_vector / compute_step,
// src/optimization/linearOptimizer.cpp:330-505) from the region fields of one
// candidate batch: candidate 0 at x, candidate i + 1 at x + dx e_i.  csim stacks the
// real and imaginary parts of (E(x + dx e_i) - E(x)) / dx over every slice into the
// Jacobian's column i and -E(x) into the objective vector; the least-squares step
// with regularisation mu solves (J^T J + mu^2 I) d = J^T e, where
//   (J^T J)(i, j) = sum_q Re(conj(D_i(q)) D_j(q)),  (J^T e)(i) = sum_q Re(conj(D_i(q)) e(q)),
// q running over (slice, region pixel), D_i = (F_{i+1} - F_0) / dx, e = -F_0.
#include "common.hpp"
#include "linopt.hpp"

namespace csx {

// F: nslice x ncand x npix complex (slice-major).  One 16 x 16 block of the Gram matrix
// per workgroup (256 threads); the q range goes through LDS in chunks of 64.
constexpr int kGramTile = 16, kGramChunk = 64;

__global__ __launch_bounds__(256) void k_lin_gram(const double2* __restrict__ F, int nslice, int ncand, int npix,
                                                  double inv_dx, double* __restrict__ A, int n) {
  __shared__ double2 Di[kGramChunk][kGramTile + 1], Dj[kGramChunk][kGramTile + 1];
  const int i0 = blockIdx.y * kGramTile, j0 = blockIdx.x * kGramTile;
  const int ti = threadIdx.x / kGramTile, tj = threadIdx.x % kGramTile;
  const long Q = (long)nslice * npix;
  double acc = 0.0;
  for (long q0 = 0; q0 < Q; q0 += kGramChunk) {
    // load D for the tile's 16 i and 16 j columns over 64 q values: 2 x 1024 values
    for (int e = threadIdx.x; e < kGramChunk * kGramTile; e += blockDim.x) {
      const int qq = e / kGramTile, c = e % kGramTile;
      const long q = q0 + qq;
      double2 di = make_double2(0.0, 0.0), dj = make_double2(0.0, 0.0);
      if (q < Q) {
        const long s = q / npix, p = q - s * npix;
        const double2* Fs = F + s * (long)ncand * npix + p;
        const double2 f0 = Fs[0];
        if (i0 + c < n) {
          const double2 f = Fs[(long)(i0 + c + 1) * npix];
          di = make_double2((f.x - f0.x) * inv_dx, (f.y - f0.y) * inv_dx);
        }
        if (j0 + c < n) {
          const double2 f = Fs[(long)(j0 + c + 1) * npix];
          dj = make_double2((f.x - f0.x) * inv_dx, (f.y - f0.y) * inv_dx);
        }
      }
      Di[qq][c] = di;
      Dj[qq][c] = dj;
    }
    __syncthreads();
#pragma unroll 8
    for (int qq = 0; qq < kGramChunk; qq++) {
      const double2 a = Di[qq][ti], b = Dj[qq][tj];
      acc += a.x * b.x + a.y * b.y;
    }
    __syncthreads();
  }
  if (i0 + ti < n && j0 + tj < n) A[(long)(i0 + ti) * n + j0 + tj] = acc;
}

// b(i) = sum_q Re(conj(D_i(q)) e(q)), e = -F_0: one workgroup per i.
__global__ __launch_bounds__(256) void k_lin_rhs(const double2* __restrict__ F, int nslice, int ncand, int npix,
                                                 double inv_dx, double* __restrict__ b) {
  __shared__ double red[4];
  const int i = blockIdx.x;
  const long Q = (long)nslice * npix;
  double acc = 0.0;
  for (long q = threadIdx.x; q < Q; q += blockDim.x) {
    const long s = q / npix, p = q - s * npix;
    const double2* Fs = F + s * (long)ncand * npix + p;
    const double2 f0 = Fs[0], f = Fs[(long)(i + 1) * npix];
    const double dr = (f.x - f0.x) * inv_dx, di = (f.y - f0.y) * inv_dx;
    acc += dr * -f0.x + di * -f0.y;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) b[i] = red[0] + red[1] + red[2] + red[3];
}

int linear_normal_equations(const void* d_F, int nslice, int n, int npix, double dx, double* d_A, double* d_b,
                            hipStream_t s) {
  if (n <= 0 || nslice <= 0 || npix <= 0 || dx == 0) return fail(CSX_ERR_ARG, "linear step: bad sizes");
  const int nt = (n + kGramTile - 1) / kGramTile;
  hipLaunchKernelGGL(k_lin_gram, dim3(nt, nt), dim3(256), 0, s, static_cast<const double2*>(d_F), nslice, n + 1,
                     npix, 1.0 / dx, d_A, n);
  hipLaunchKernelGGL(k_lin_rhs, dim3(n), dim3(256), 0, s, static_cast<const double2*>(d_F), nslice, n + 1, npix,
                     1.0 / dx, d_b);
  CSX_HIP(hipGetLastError());
  return CSX_OK;
}

}  // namespace csx
