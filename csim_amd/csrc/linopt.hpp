// Device side of csim's linear optimiser step (src/optimization/linearOptimizer.cpp).
#pragma once
#include "common.hpp"

namespace csx {

// From the region fields of one candidate batch (d_F: nslice x (n + 1) x npix complex,
// candidate 0 at x, candidate i + 1 at x + dx e_i) the normal equations of the
// linearised step: d_A (n x n, row-major) = J^T J and d_b (n) = J^T e, with J's column i
// the real and imaginary parts of (E(x + dx e_i) - E(x)) / dx and e = -E(x).
int linear_normal_equations(const void* d_F, int nslice, int n, int npix, double dx, double* d_A, double* d_b,
                            hipStream_t s);

}  // namespace csx
