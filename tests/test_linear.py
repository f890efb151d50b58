"""linearOptimizer restated over the batched chain (CPU: a linear stand-in chain).

The stand-in returns region fields E(x) = E0 + G x exactly, so the finite-difference
Jacobian is G, the regularised step must equal the stacked least-squares solution of
[G; mu I] d = [-E; 0] (numpy lstsq, an independent route), and the mu scan must keep
the last mu that lowered the contrast (csim's compute_step rule)."""
import numpy as np
import torch

from csim_amd.linear import LinearConfig, LinearOptimizer


class LinearChain:
    def __init__(self, E0, G, scale=1.0):
        self.E0, self.G, self.scale = E0, G, scale      # E0 (nslice, npix), G (nslice, npix, n)
        self.X = None

    def set_candidates(self, elem, X):
        self.X = np.asarray(X, dtype=np.float64)

    def _fields(self):
        return self.E0[None] + np.einsum("spn,cn->csp", self.G, self.X)

    def contrast(self):
        F = self._fields()
        return (np.abs(F) ** 2).mean(axis=(1, 2)) * self.scale, {}

    def region_fields(self, **kw):
        return torch.from_numpy(np.ascontiguousarray(np.transpose(self._fields(), (1, 0, 2))))


def make(seed=3, nslice=3, npix=40, n=6):
    rng = np.random.default_rng(seed)
    E0 = rng.standard_normal((nslice, npix)) + 1j * rng.standard_normal((nslice, npix))
    G = rng.standard_normal((nslice, npix, n)) + 1j * rng.standard_normal((nslice, npix, n))
    return LinearChain(E0, G), E0, G


def test_jacobian_and_objective_vector():
    ch, E0, G = make()
    opt = LinearOptimizer(ch, "fpm", LinearConfig(jacobian_dx=1e-3, global_bounds=(-1e9, 1e9)))
    x = np.linspace(-0.1, 0.1, 6)
    c, J, e = opt.jacobian(x)
    E = E0 + np.einsum("spn,n->sp", G, x)
    Jref = np.concatenate([G.real, G.imag], axis=1).reshape(-1, 6)      # per slice: real rows, imag rows
    assert np.allclose(J.numpy(), Jref, rtol=1e-9, atol=1e-9)
    assert np.allclose(e.numpy(), -np.concatenate([E.real, E.imag], axis=1).ravel(), rtol=1e-14)
    assert abs(c - (np.abs(E) ** 2).mean()) <= 1e-14 * c


def test_step_is_stacked_least_squares_and_mu_rule():
    ch, E0, G = make(seed=5)
    cfg = LinearConfig(jacobian_dx=1e-3, mu_scan=(0.5, 8.0, 0.5), global_bounds=(-1e9, 1e9))
    opt = LinearOptimizer(ch, "fpm", cfg)
    x = np.zeros(6)
    c0, J, e = opt.jacobian(x)
    d, mu, cbest = opt.step(x, J, e)
    Jn, en = J.numpy(), e.numpy()
    mus = np.arange(0.5, 8.0 + 0.25, 0.5)
    cs = []
    for m in mus:                                                  # independent route: stacked lstsq
        Aug = np.vstack([Jn, m * np.eye(6)])
        rhs = np.concatenate([en, np.zeros(6)])
        dm = np.linalg.lstsq(Aug, rhs, rcond=None)[0]
        E = E0 + np.einsum("spn,n->sp", G, x + dm)
        cs.append(((np.abs(E) ** 2).mean(), dm))
    k = 0
    while k + 1 < len(cs) and cs[k + 1][0] < cs[k][0]:
        k += 1
    assert mu == mus[k]
    assert np.allclose(d, cs[k][1], rtol=1e-9, atol=1e-12)
    assert abs(cbest - cs[k][0]) <= 1e-12 * cs[k][0]


def test_run_lowers_contrast_and_respects_bounds():
    ch, E0, G = make(seed=7)
    cfg = LinearConfig(jacobian_dx=1e-4, mu_scan=(0.1, 3.0, 0.1), global_bounds=(-0.5, 0.5),
                       max_iterations=5, contrast_rel_tolerance=0.0, stop_contrast=0.0)
    res = LinearOptimizer(ch, "fpm", cfg).run(np.zeros(6))
    cs = [h[2] for h in res.history]
    assert res.best_value < cs[0]
    assert np.all(np.abs(res.best_vector) <= 0.5)
