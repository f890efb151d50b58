"""linearOptimizer (src/optimization/linearOptimizer.cpp) over the batched GPU chain.

csim's regularised linear (field-conjugation) optimiser of one optimisation vector
(here the sags of a hex FPM):

  compute_jacobian      :420-488  column i = (E(x + dx e_i) - E(x)) / dx at the region
                                  pixels, real parts then imaginary parts per slice
  compute_objective_vector :490-505  -E(x) in the same row order
  compute_step          :330-405  [J; mu I] d = [-E; 0] for mu = 1e7, 2e7, ... 3e8
                                  while the contrast keeps decreasing; the last mu
                                  that decreased it is kept
  optimize              :228-300  x += d, clipped to the global bounds, until a
                                  stopping rule (:302-313) holds

Here the n + 1 propagations of the Jacobian are one candidate batch, the whole mu
scan is one more batch (its contrasts are then read in csim's order), and every
least-squares step is solved on the GPU in FP64 from the normal equations
(J^T J + mu^2 I) d = J^T (-E), which have the same solution as the stacked system.
The stopping rule on the contrast compares an absolute difference with
`contrast_rel_tolerance` starting from a previous contrast of 0, as csim does.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class LinearConfig:
    jacobian_dx: float = 1e-8                 # jacobianDx
    mu_scan: tuple = (1e7, 3e8, 1e7)          # compute_step's regularisation scan (start, stop, step)
    global_bounds: tuple = (0.0, 1e-6)        # globalBounds
    max_iterations: int = 20                  # maxIterations
    stop_contrast: float = 1e-9               # stopContrast
    contrast_rel_tolerance: float = 1e-5      # contrastRelTolerance (an absolute difference)
    delta_tolerance: float = 1e-9             # deltaOptVecTolerance


@dataclass
class LinearResult:
    best_value: float
    best_vector: np.ndarray
    history: list = field(default_factory=list)     # (iteration, wall seconds, contrast, mu)
    propagations: int = 0
    seconds: float = 0.0


class LinearOptimizer:
    """Optimises the sag vector of the hex-FPM element `element` of a chain (one
    compiled chain, its candidate batch resized as needed)."""

    def __init__(self, chain, element, cfg: LinearConfig = LinearConfig(), region=None):
        self.ch, self.elem, self.c = chain, element, cfg
        self.region = region or {}
        self.props = 0

    def _contrasts(self, X):
        """Contrast of every row of X (candidates x n) in one candidate batch."""
        self.ch.set_candidates(self.elem, np.ascontiguousarray(X))
        c, _ = self.ch.contrast()
        self.props += X.shape[0]
        return np.asarray(c)

    def _real_rows(self, F):
        """(nslice, ncand, npix) complex -> (ncand, 2 nslice npix) real, per slice the
        real parts then the imaginary parts (compute_jacobian's row order)."""
        import torch
        return torch.cat([F.real, F.imag], dim=2).permute(1, 0, 2).reshape(F.shape[1], -1)

    def jacobian(self, x):
        """Returns (contrast at x, J (rows x n), e = -E(x) (rows)) as FP64 device tensors."""
        n, dx = x.size, self.c.jacobian_dx
        X = np.repeat(x[None, :], n + 1, axis=0)
        X[1:] += dx * np.eye(n)
        c = self._contrasts(X)
        R = self._real_rows(self.ch.region_fields(**self.region))
        J = ((R[1:] - R[:1]) / dx).T.contiguous()
        return float(c[0]), J, -R[0]

    def step(self, x, J, e):
        """compute_step: the regularised step for each mu of the scan, their contrasts
        in one batch, the last mu that kept lowering the contrast."""
        import torch
        A = J.T @ J
        b = J.T @ e
        n = x.size
        I = torch.eye(n, dtype=torch.float64, device=A.device)
        mus = np.arange(self.c.mu_scan[0], self.c.mu_scan[1] + 0.5 * self.c.mu_scan[2], self.c.mu_scan[2])
        D = torch.stack([torch.linalg.solve(A + (mu * mu) * I, b) for mu in mus]).cpu().numpy()
        cs = self._contrasts(x[None, :] + D)
        last, k_best = 1e6, 0
        for k, ck in enumerate(cs):
            if ck < last:
                last, k_best = ck, k
            else:
                break
        return D[k_best], float(mus[k_best]), float(last)

    def run(self, x0):
        c = self.c
        t0 = time.perf_counter()
        x = np.clip(np.asarray(x0, dtype=np.float64).copy(), *c.global_bounds)
        best, best_x, hist = np.inf, x.copy(), []
        last_contrast, delta_norm, mu, it = 0.0, np.inf, None, 0
        while True:
            contrast, J, e = self.jacobian(x)
            hist.append((it, time.perf_counter() - t0, contrast, mu))
            if contrast < best:
                best, best_x = contrast, x.copy()
            stop = (abs(last_contrast - contrast) < c.contrast_rel_tolerance or contrast < c.stop_contrast or
                    delta_norm < c.delta_tolerance or it > c.max_iterations)
            last_contrast = contrast
            if stop:
                break
            d, mu, _ = self.step(x, J, e)
            delta_norm = float(np.linalg.norm(d))
            x = np.clip(x + d, *c.global_bounds)
            it += 1
        return LinearResult(best, best_x, hist, self.props, time.perf_counter() - t0)
