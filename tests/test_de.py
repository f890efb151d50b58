"""DE driver (host logic) on the CPU: csim's update rules and convergence."""
import os

import numpy as np

import oracle
from csim_amd.de import DEConfig, DifferentialEvolution, population_permutations

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_permutations_are_distinct_rotations():
    rng = np.random.Generator(np.random.Philox(1))
    P = 11
    pop = np.arange(P, dtype=float)[None, :].repeat(3, axis=0)
    perms = population_permutations(pop, 3, rng)
    ids = np.stack([p[0] for p in perms])          # member index used in each slice
    for col in range(P):
        assert len(set(ids[:, col])) == 3              # three different members per candidate
    for p in perms:
        assert sorted(p[0]) == list(range(P))          # each slice is a permutation


def test_bounce_back_respects_bounds_and_converges_on_quadratic():
    D = 6
    target = np.linspace(0.1, 0.9, D)
    calls = []

    def f(pop):
        calls.append(pop.copy())
        return np.sum((pop - target[:, None]) ** 2, axis=0)

    cfg = DEConfig(population_size=40, max_iterations=300, global_bounds=(0.0, 1.0), init_bounds=(-0.5, 0.5),
                   init_mean=np.full(D, 0.5), output_interval=10)
    res = DifferentialEvolution(f, D, cfg).run()
    for pop in calls:
        assert pop.min() >= -1e-12 and pop.max() <= 1 + 1e-12
    assert res.best_value < 1e-6
    assert np.allclose(res.best_vector, target, atol=2e-3)


def test_de_on_reference_response_fixture():
    """A short DE over csim's shipped linear response (mono 650 nm, 499 hexes) improves
    the dark-hole objective from the near-zero-sag start by > 10x."""
    fx = np.load(os.path.join(GOLDEN, "mono_kat.npz"))
    M = fx["M"][:, :, None]
    lam, calib = float(fx["wavelength"]), float(fx["calib"])

    def f(pop):
        return oracle.hex_resp_objective(M, [lam], 1.0, calib, pop)

    cfg = DEConfig(population_size=100, max_iterations=150, global_bounds=(-5e-7, 5e-7),
                   init_bounds=(-1e-7, 1e-7), init_mean=np.zeros(499), start_strategy="deLocalToBest1",
                   output_interval=50)
    res = DifferentialEvolution(f, 499, cfg).run()
    start = f(np.zeros((499, 1)))[0]
    assert res.best_value < start / 10
    assert res.best_value > 1.3e-10 * 0.5        # cannot beat the reference's long-run optimum by much


def test_device_population_same_rules():
    """The population held as torch tensors (here on the CPU device) follows the same
    rules: every candidate inside the bounds, convergence on the quadratic, and a
    screen that evaluates a subset sees tensors and changes only those members."""
    import torch
    D = 6
    target = np.linspace(0.1, 0.9, D)
    seen = []

    def f(pop):
        assert torch.is_tensor(pop) and pop.dtype == torch.float64
        seen.append((float(pop.min()), float(pop.max())))
        return ((pop - torch.as_tensor(target)[:, None]) ** 2).sum(dim=0)

    cfg = DEConfig(population_size=40, max_iterations=300, global_bounds=(0.0, 1.0), init_bounds=(-0.5, 0.5),
                   init_mean=np.full(D, 0.5), output_interval=10)
    res = DifferentialEvolution(f, D, cfg, device="cpu").run()
    assert min(lo for lo, _ in seen) >= -1e-12 and max(hi for _, hi in seen) <= 1 + 1e-12
    assert res.best_value < 1e-6
    assert np.allclose(res.best_vector, target, atol=2e-3)
    assert isinstance(res.population, np.ndarray) and res.population.shape == (D, 40)

    picked = []

    def screen(cand):
        assert torch.is_tensor(cand)
        idx = np.arange(0, cand.shape[1], 4)
        picked.append(idx)
        return idx, f(cand[:, torch.as_tensor(idx)])

    cfg2 = DEConfig(population_size=40, max_iterations=0, global_bounds=(0.0, 1.0), init_bounds=(-0.5, 0.5),
                    init_mean=np.full(D, 0.5))                # one generation
    de = DifferentialEvolution(f, D, cfg2, screen=screen, device="cpu")
    init = de.init_population(de.ops.vec(np.full(D, 0.5)))
    res2 = de.run(init_pop=init, init_val=f(init))
    assert len(picked) == 1 and res2.evaluations == 10
    untouched = np.setdiff1d(np.arange(40), picked[0])
    assert np.array_equal(res2.population[:, untouched], init.numpy()[:, untouched])
    assert not np.array_equal(res2.population[:, picked[0]], init.numpy()[:, picked[0]])
