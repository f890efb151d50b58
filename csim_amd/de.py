"""Differential evolution over hex-FPM sags, evaluated in candidate batches.

Restates csim's optimiser (src/optimization/differentialEvolution/):
  strategies          deStrategy.cpp:23-158 (DE/rand/1, DE/local-to-best/1, either-or),
                      population permutations by successive rotations of one shuffle
  crossover           differentialEvolutionOptimizer.cpp:360-369
  bounce-back bounds  :395-434 (around the population mean, clipped to the global bounds)
  selection           :437-455
  restarts/switching  :216-330 (strategy switches on population spread, small-change reinit)
The objective is any callable mapping a (D, P) population to P values; the
batched GPU chain evaluates every candidate of a generation in one pass.
Random numbers come from numpy's Philox generator (the reference's arma/rand()
streams are not reproduced; trajectories differ, the update rules do not).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


def _rotations(P, n, rng, perm_P):
    """Index arrays of makePopulationPermutations: slice i takes the members si1_i,
    si1_{i+1} = si1_i rotated by si2[i] + 1 (si1 from perm_P, si2 from rng)."""
    si2 = rng.permutation(n)
    si1 = perm_P(P)
    out = []
    for i in range(n):
        out.append(si1)
        if i < n - 1:
            k = int(si2[i]) + 1
            si1 = si1[(np.arange(P) - k) % P] if isinstance(si1, np.ndarray) else si1.roll(k)
    return out


def population_permutations(pop, n, rng):
    """makePopulationPermutations: slice i = pop[:, si1_i], si1_{i+1} = si1_i rotated by si2[i]+1."""
    return [pop[:, ix] for ix in _rotations(pop.shape[1], n, rng, rng.permutation)]


class _NumpyOps:
    """Array operations of the DE on the host (numpy, Philox stream)."""

    def __init__(self, seed):
        self.rng = np.random.Generator(np.random.Philox(seed))

    def vec(self, x):
        return np.asarray(x, dtype=np.float64)

    def rand(self, shape):
        return self.rng.random(shape)

    def coin(self):
        return self.rng.random()

    def perms(self, pop, n):
        return [pop[:, ix] for ix in _rotations(pop.shape[1], n, self.rng, self.rng.permutation)]

    where, minimum, maximum = staticmethod(np.where), staticmethod(np.minimum), staticmethod(np.maximum)

    def spread(self, pop):
        mean = pop.mean(axis=1)
        d = np.linalg.norm(pop - mean[:, None], axis=0)
        return mean, float(d.mean()), float(d.max())

    def argmin(self, val):
        i = int(np.argmin(val))
        return i, float(val[i])

    def full(self, n, v):
        return np.full(n, v, dtype=np.float64)

    def index(self, idx):
        return np.asarray(idx, dtype=int)

    def host(self, x):
        return np.asarray(x)

    def copy(self, x):
        return np.array(x, dtype=np.float64)


class _TorchOps:
    """The same operations on a torch device: a generation of a large population is a
    few kernels instead of host passes over it.  Bulk random numbers come from a device
    generator, the few small integers and coin flips of the strategies from a host
    Philox stream."""

    def __init__(self, seed, device):
        import torch
        self.t, self.dev = torch, torch.device(device)
        self.gen = torch.Generator(device=self.dev)
        self.gen.manual_seed(seed)
        self.rng = np.random.Generator(np.random.Philox(seed))

    def vec(self, x):
        return self.t.as_tensor(np.asarray(x, dtype=np.float64) if not self.t.is_tensor(x) else x,
                                dtype=self.t.float64, device=self.dev)

    def rand(self, shape):
        return self.t.rand(shape, generator=self.gen, dtype=self.t.float64, device=self.dev)

    def coin(self):
        return self.rng.random()

    def perms(self, pop, n):
        perm_P = lambda P: self.t.randperm(P, generator=self.gen, device=self.dev)
        return [pop[:, ix] for ix in _rotations(pop.shape[1], n, self.rng, perm_P)]

    def where(self, m, a, b):
        return self.t.where(m, a, b)

    def minimum(self, a, b):
        return self.t.minimum(a, b)

    def maximum(self, a, b):
        return self.t.maximum(a, b)

    def spread(self, pop):
        mean = pop.mean(dim=1)
        d = (pop - mean[:, None]).norm(dim=0)
        st = self.t.stack([d.mean(), d.max()]).cpu()
        return mean, float(st[0]), float(st[1])

    def argmin(self, val):
        i = int(self.t.argmin(val))
        return i, float(val[i])

    def full(self, n, v):
        return self.t.full((n,), v, dtype=self.t.float64, device=self.dev)

    def index(self, idx):
        return self.t.as_tensor(np.asarray(idx, dtype=np.int64), device=self.dev)

    def host(self, x):
        return x.cpu().numpy()

    def copy(self, x):
        return x.clone()


def de_rand1(pop, best, F, ops):
    p = ops.perms(pop, 3)
    return p[2] + F * (p[0] - p[1]), p[2]


def de_local_to_best1(pop, best, F, ops):
    p = ops.perms(pop, 2)
    return pop + F * (best[:, None] - pop) + F * (p[0] - p[1]), pop


def either_or(pop, best, F, ops):
    p = ops.perms(pop, 3)
    if ops.coin() < 0.5:
        return p[2] + F * (p[0] - p[1]), p[2]
    return p[2] + 0.5 * (F + 1.0) * (p[0] + p[1] - 2 * p[2]), p[2]


STRATEGIES = {"deRand1": de_rand1, "deLocalToBest1": de_local_to_best1, "eitherOr": either_or}


@dataclass
class DEConfig:
    population_size: int = 64
    start_strategy: str = "deLocalToBest1"
    base_strategy: str = "eitherOr"
    end_strategy: str = "deLocalToBest1"
    reinit_strategy: str = "eitherOr"
    f_weight: float = 0.5
    crossover_prob: float = 1.0
    global_bounds: tuple = (0.0, 1e-6)
    init_bounds: tuple = (0.0, 1e-6)
    init_mean: np.ndarray | None = None
    init_spread: float = 1.0
    reinit_compression: float = 0.1
    reinit_spread: float = 1e-12
    max_iterations: int = 100
    stop_objective_value: float = 0.0
    stop_population_spread: float = 0.0
    reinit_changefrac_threshold: float = 0.001
    max_small_change: int = 5
    max_retries: int = 1000
    base_strategy_spread_switch: float = 0.0
    end_strategy_spread_switch: float = 0.0
    output_interval: int = 10
    seed: int = 20261206


@dataclass
class DEResult:
    best_value: float
    best_vector: np.ndarray
    history: list = field(default_factory=list)     # (iteration, wall seconds, best value, evaluations)
    evaluations: int = 0
    seconds: float = 0.0
    population: np.ndarray | None = None            # final population (D, P) and its values
    values: np.ndarray | None = None


class DifferentialEvolution:
    def __init__(self, objective, dim, cfg: DEConfig, screen=None, device=None):
        """objective(pop (D, P)) -> values (P).  screen(pop (D, K)) -> indices of the
        candidates worth evaluating exactly (surrogate screening), or (indices, exact
        values) when the screen evaluates its survivors itself; None evaluates all.
        device None keeps the population in numpy arrays; a torch device ("cuda") keeps
        it there as float64 tensors, which objective and screen then receive."""
        self.f = objective
        self.D = dim
        self.c = cfg
        self.screen = screen
        self.ops = _NumpyOps(cfg.seed) if device is None else _TorchOps(cfg.seed, device)
        ops = self.ops
        gl, gu = cfg.global_bounds
        il, iu = cfg.init_bounds
        self.glo, self.ghi = ops.full(dim, gl), ops.full(dim, gu)
        self.ilo, self.ihi = ops.full(dim, il), ops.full(dim, iu)
        self.evals = 0

    def _eval(self, pop):
        self.evals += pop.shape[1]
        return self.ops.vec(self.f(pop))

    def init_population(self, origin, spread=1.0):
        c = self.c
        P = c.population_size
        return origin[:, None] + spread * c.init_spread * (self.ops.rand((self.D, P)) - 0.5) * \
            (self.ihi - self.ilo)[:, None]

    def run(self, time_budget=None, target=None, eval_budget=None, init_pop=None, init_val=None,
            stop=None) -> DEResult:
        """Iterates until the csim stopping rules, a wall-time budget (s), a target
        objective, a budget of exact evaluations or stop() is reached.  init_pop (D, P)
        (with its values init_val, or evaluated here) replaces the initial population."""
        c, ops = self.c, self.ops
        t0 = time.perf_counter()
        if init_pop is not None:
            pop = ops.copy(ops.vec(init_pop))
            val = ops.vec(init_val) if init_val is not None else self._eval(pop)
        else:
            origin = ops.vec(c.init_mean) if c.init_mean is not None else 0.5 * (self.ilo + self.ihi)
            pop = self.init_population(origin)
            val = self._eval(pop)
        strategy = STRATEGIES[c.start_strategy]
        best_val, best_vec = np.inf, None
        gbest_val, gbest_vec = np.inf, None
        hist = []

        def find_best():
            nonlocal best_val, best_vec, gbest_val, gbest_vec
            i, v = ops.argmin(val)
            if v < best_val:
                best_val, best_vec = v, ops.copy(pop[:, i])
                if best_val < gbest_val:
                    gbest_val, gbest_vec = best_val, best_vec

        find_best()
        mean, spread_mean, spread_max = ops.spread(pop)
        hist.append((0, time.perf_counter() - t0, gbest_val, self.evals))
        small, prev_best, retries = 0, 0.0, 0
        sel_base = sel_end = False
        do_reinit = False
        it = 0
        while (it == 0 or not (it > c.max_iterations or best_val < c.stop_objective_value or
                               spread_max < c.stop_population_spread)) and retries <= c.max_retries:
            if time_budget is not None and time.perf_counter() - t0 > time_budget:
                break
            if target is not None and gbest_val <= target:
                break
            if eval_budget is not None and self.evals >= eval_budget:
                break
            if stop is not None and stop():
                break
            if small >= c.max_small_change or do_reinit or spread_mean < c.reinit_spread:
                best_val = np.inf
                strategy = STRATEGIES[c.reinit_strategy]
                pop = self.init_population(gbest_vec, c.reinit_compression)
                val = self._eval(pop)
                find_best()
                small, do_reinit = 0, False
                retries += 1
            mutated, original = strategy(pop, best_vec, c.f_weight, ops)
            if c.crossover_prob < 1.0:
                cand = ops.where(ops.rand(tuple(pop.shape)) < c.crossover_prob, mutated, pop)
            else:
                cand = mutated
            # bounce-back within [mean + init bounds] clipped to the global bounds
            up = ops.minimum(mean + self.ihi, self.ghi)
            lo = ops.maximum(mean + self.ilo, self.glo)
            r = ops.rand(tuple(cand.shape))
            below = cand < lo[:, None]
            above = ~below & (cand > up[:, None])
            cand = ops.where(below, lo[:, None] + r * (original - lo[:, None]), cand)
            cand = ops.where(above, up[:, None] + r * (original - up[:, None]), cand)
            # selection (optionally on a surrogate-screened subset)
            if self.screen is not None:
                res = self.screen(cand)
                cv = ops.full(cand.shape[1], np.inf)
                if isinstance(res, tuple):          # the screen evaluated its survivors itself
                    idx = np.asarray(res[0], dtype=int)
                    self.evals += idx.size
                    cv[ops.index(idx)] = ops.vec(res[1])
                else:
                    idx = np.asarray(res, dtype=int)
                    if idx.size:
                        cv[ops.index(idx)] = self._eval(cand[:, ops.index(idx)])
            else:
                cv = self._eval(cand)
            better = cv < val
            pop = ops.where(better[None, :], cand, pop)
            val = ops.where(better, cv, val)
            mean, spread_mean, spread_max = ops.spread(pop)
            find_best()
            if it % c.output_interval == 0:
                if spread_mean < c.base_strategy_spread_switch and not sel_base:
                    strategy, sel_base, do_reinit = STRATEGIES[c.base_strategy], True, True
                if spread_mean < c.end_strategy_spread_switch and not sel_end:
                    strategy, sel_end, do_reinit = STRATEGIES[c.end_strategy], True, True
                if best_val != 0 and abs((best_val - prev_best) / best_val) < c.reinit_changefrac_threshold:
                    small += 1
                else:
                    small = 0
                prev_best = best_val
            it += 1
            hist.append((it, time.perf_counter() - t0, gbest_val, self.evals))
        return DEResult(gbest_val, ops.host(gbest_vec), hist, self.evals, time.perf_counter() - t0, ops.host(pop),
                        ops.host(val))
