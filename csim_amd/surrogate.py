"""Learned surrogate for DE screening (no reference counterpart).

An MLP  (n_hex thicknesses + wavelength) -> 32 annulus log10-contrasts over
3-10 lambda/D (thicknesses embedded as hex transmission phasors), trained on propagations of the batched GPU chain with the
library's own kernels (csx_mlp_*: BF16 operands, FP32 accumulation, Adam).
The DE objective (dark-hole mean of sum_lambda I / peak) is rebuilt from the
predicted per-wavelength annulus contrasts; the surrogate ranks a large
candidate population and the exact propagation re-evaluates the best few.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

N_BINS = 32
R_IN, R_OUT = 3.0, 10.0
T_SCALE = 1e-6          # thickness normalisation (1 um)
LAM_CENTER, LAM_SCALE = 630e-9, 30e-9
LOG_FLOOR = 1e-14


def annulus_bins(chain, n_bins=N_BINS, r_in=R_IN, r_out=R_OUT):
    """Bin index (or -1) of each output pixel: equal-width annuli in lambda/D."""
    inf = chain.info()
    n, d, loD = inf["out_rows"], inf["out_pixel_size"], inf["lambda_over_d"]
    x = -(n * d - d) / 2 + np.arange(n) * d
    r = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2) / loD
    b = np.floor((r - r_in) / (r_out - r_in) * n_bins).astype(np.int32)
    b[(r < r_in) | (r > r_out)] = -1
    b[b >= n_bins] = n_bins - 1
    counts = np.bincount(b[b >= 0].ravel(), minlength=n_bins)
    return b, counts


def labels_from_chain(chain, bins, counts, peak):
    """Per (candidate, lambda) annulus mean contrast (sources summed) -> (n_cand, n_lambda, n_bins)."""
    inf = chain.info()
    prof = chain.field_profiles(bins, len(counts))
    prof = prof.reshape(inf["n_candidates"], inf["n_sources"], inf["n_lambdas"], len(counts)).sum(axis=1)
    return prof / counts[None, None, :] / peak


N_INDEX = 1.46          # hex material index: phase 2 pi (n - 1) t / lambda


def features(thickness, lambdas):
    """(n_cand, n_hex) thicknesses x n_lambda -> (n_cand * n_lambda, 2 n_hex + 1) float32.

    The network's input is the thickness vector and the wavelength; its first,
    parameter-free layer is the physical embedding of each hex as its transmission
    phasor cos / sin(2 pi (n-1) t / lambda), in which the focal-plane field is linear
    (and the annulus intensities quadratic), followed by the normalised wavelength."""
    t = np.asarray(thickness, dtype=np.float64)
    lam = np.asarray(lambdas, dtype=np.float64)
    nc, nh = t.shape
    nl = lam.size
    ph = 2 * np.pi * (N_INDEX - 1.0) * t[:, None, :] / lam[None, :, None]     # (nc, nl, nh)
    X = np.empty((nc, nl, 2 * nh + 1), dtype=np.float32)
    X[:, :, :nh] = np.cos(ph)
    X[:, :, nh:2 * nh] = np.sin(ph)
    X[:, :, 2 * nh] = (lam - LAM_CENTER) / LAM_SCALE
    return X.reshape(nc * nl, 2 * nh + 1)


def features_device(ctx, thickness, lambdas):
    """`features` computed on the GPU (csx_hex_phasor_features) -> float32 device tensor."""
    import torch
    if torch.is_tensor(thickness):
        t = thickness.to(device="cuda", dtype=torch.float64).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(thickness, dtype=np.float64)).cuda()
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    nc, nh = t.shape
    out = torch.empty((nc * lam.size, 2 * nh + 1), dtype=torch.float32, device=t.device)
    check(lib().csx_hex_phasor_features(ctx.handle, C.c_void_p(t.data_ptr()), nc, nh,
                                        lam.ctypes.data_as(C.c_void_p), lam.size, N_INDEX, LAM_CENTER,
                                        LAM_SCALE, C.c_void_p(out.data_ptr())))
    return out


def objective_from_profiles(C_lam_bins, counts):
    """Dark-hole mean of sum_lambda I / peak from per-lambda annulus means."""
    w = counts / counts.sum()
    return np.einsum("clb,b->c", C_lam_bins, w)


class MLP:
    """csx_mlp handle: n_in -> h1 -> h2 -> n_out GELU perceptron on the GPU."""

    def __init__(self, ctx, n_in, n_h1=512, n_h2=512, n_out=N_BINS, seed=1):
        self.ctx = ctx
        p = C.c_void_p()
        check(lib().csx_mlp_create(ctx.handle, n_in, n_h1, n_h2, n_out, seed, C.byref(p)))
        self._p = p
        self.n_in, self.n_out = n_in, n_out
        ctx.adopt(self)

    def close(self):
        if self._p:
            lib().csx_mlp_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        n = lib().csx_mlp_param_count(self._p)
        flat = np.zeros(n)
        check(lib().csx_mlp_get_params(self._p, flat.ctypes.data_as(C.c_void_p)))
        return flat

    def set_params(self, flat):
        """Replaces the parameters (the Adam state restarts)."""
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        check(lib().csx_mlp_set_params(self._p, flat.ctypes.data_as(C.c_void_p)))

    def train_step(self, x_dev, t_dev, lr, want_loss=True):
        loss = C.c_double()
        check(lib().csx_mlp_train_step(self._p, C.c_void_p(x_dev.data_ptr()), C.c_void_p(t_dev.data_ptr()),
                                       x_dev.shape[0], float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def forward(self, x_dev):
        import torch
        y = torch.empty((x_dev.shape[0], self.n_out), dtype=torch.float32, device=x_dev.device)
        check(lib().csx_mlp_forward(self._p, C.c_void_p(x_dev.data_ptr()), x_dev.shape[0],
                                    C.c_void_p(y.data_ptr())))
        return y

    def fit(self, X, Y, epochs=20, batch=1024, lr=1e-3, seed=0, log=None):
        """X (n, n_in), Y (n, n_out) float32 (numpy or device tensors), targets log10 contrast."""
        import torch
        rng = np.random.Generator(np.random.Philox(seed))
        Xd = X if torch.is_tensor(X) else torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32))
        Yd = Y if torch.is_tensor(Y) else torch.from_numpy(np.ascontiguousarray(Y, dtype=np.float32))
        Xd, Yd = Xd.cuda().contiguous(), Yd.cuda().contiguous()
        n = X.shape[0]
        hist = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).cuda()
            tot = 0.0
            nb = 0
            for i in range(0, n - batch + 1, batch):
                idx = perm[i:i + batch]
                xb = Xd.index_select(0, idx).contiguous()
                yb = Yd.index_select(0, idx).contiguous()
                tot += self.train_step(xb, yb, lr)
                nb += 1
            hist.append(tot / max(1, nb))
            if log:
                log(ep, hist[-1])
        return hist

    def predict(self, X):
        import torch
        Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
        return self.forward(Xd).cpu().numpy()


class Surrogate:
    """Screens DE candidates: predicted dark-hole objective from the MLP."""

    def __init__(self, mlp: MLP, lambdas, counts):
        self.mlp = mlp
        self.lambdas = list(lambdas)
        self.counts = np.asarray(counts, dtype=np.float64)

    def predict_objective(self, thickness):
        nc = thickness.shape[0]
        y = self.mlp.forward(features_device(self.mlp.ctx, thickness, self.lambdas)).cpu().numpy().astype(np.float64)
        C = 10.0 ** y.reshape(nc, len(self.lambdas), -1)
        return objective_from_profiles(C, self.counts)
