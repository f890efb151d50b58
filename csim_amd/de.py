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


def population_permutations(pop, n, rng):
    """makePopulationPermutations: slice i = pop[:, si1_i], si1_{i+1} = si1_i rotated by si2[i]+1."""
    P = pop.shape[1]
    si2 = rng.permutation(n)
    si1 = rng.permutation(P)
    ii1 = np.arange(P)
    out = []
    for i in range(n):
        out.append(pop[:, si1])
        if i < n - 1:
            si1 = si1[np.roll(ii1, si2[i] + 1)]
    return out


def de_rand1(pop, best, F, rng):
    p = population_permutations(pop, 3, rng)
    return p[2] + F * (p[0] - p[1]), p[2]


def de_local_to_best1(pop, best, F, rng):
    p = population_permutations(pop, 2, rng)
    return pop + F * (best[:, None] - pop) + F * (p[0] - p[1]), pop


def either_or(pop, best, F, rng):
    p = population_permutations(pop, 3, rng)
    if rng.random() < 0.5:
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
    def __init__(self, objective, dim, cfg: DEConfig, screen=None):
        """objective(pop (D, P)) -> values (P).  screen(pop (D, K)) -> indices of the
        candidates worth evaluating exactly (surrogate screening), or (indices, exact
        values) when the screen evaluates its survivors itself; None evaluates all."""
        self.f = objective
        self.D = dim
        self.c = cfg
        self.screen = screen
        self.rng = np.random.Generator(np.random.Philox(cfg.seed))
        gl, gu = cfg.global_bounds
        il, iu = cfg.init_bounds
        self.glo, self.ghi = np.full(dim, gl), np.full(dim, gu)
        self.ilo, self.ihi = np.full(dim, il), np.full(dim, iu)
        self.evals = 0

    def _eval(self, pop):
        self.evals += pop.shape[1]
        return np.asarray(self.f(pop), dtype=np.float64)

    def init_population(self, origin, spread=1.0):
        c = self.c
        P = c.population_size
        return origin[:, None] + spread * c.init_spread * (self.rng.random((self.D, P)) - 0.5) * \
            (self.ihi - self.ilo)[:, None]

    def run(self, time_budget=None, target=None, eval_budget=None, init_pop=None, init_val=None,
            stop=None) -> DEResult:
        """Iterates until the csim stopping rules, a wall-time budget (s), a target
        objective, a budget of exact evaluations or stop() is reached.  init_pop (D, P)
        (with its values init_val, or evaluated here) replaces the initial population."""
        c = self.c
        t0 = time.perf_counter()
        if init_pop is not None:
            pop = np.array(init_pop, dtype=np.float64)
            val = np.array(init_val, dtype=np.float64) if init_val is not None else self._eval(pop)
        else:
            origin = c.init_mean if c.init_mean is not None else 0.5 * (self.ilo + self.ihi)
            pop = self.init_population(origin)
            val = self._eval(pop)
        strategy = STRATEGIES[c.start_strategy]
        best_val, best_vec = np.inf, None
        gbest_val, gbest_vec = np.inf, None
        hist = []

        def find_best():
            nonlocal best_val, best_vec, gbest_val, gbest_vec
            i = int(np.argmin(val))
            if val[i] < best_val:
                best_val, best_vec = float(val[i]), pop[:, i].copy()
                if best_val < gbest_val:
                    gbest_val, gbest_vec = best_val, best_vec.copy()

        def stats():
            mean = pop.mean(axis=1)
            d = np.linalg.norm(pop - mean[:, None], axis=0)
            return mean, d.mean(), d.max()

        find_best()
        mean, spread_mean, spread_max = stats()
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
            mutated, original = strategy(pop, best_vec, c.f_weight, self.rng)
            if c.crossover_prob < 1.0:
                cross = self.rng.random(pop.shape) < c.crossover_prob
                cand = np.where(cross, mutated, pop)
            else:
                cand = mutated
            # bounce-back within [mean + init bounds] clipped to the global bounds
            up = np.minimum(mean + self.ihi, self.ghi)
            lo = np.maximum(mean + self.ilo, self.glo)
            r = self.rng.random(cand.shape)
            below = cand < lo[:, None]
            above = ~below & (cand > up[:, None])
            cand = np.where(below, lo[:, None] + r * (original - lo[:, None]), cand)
            cand = np.where(above, up[:, None] + r * (original - up[:, None]), cand)
            # selection (optionally on a surrogate-screened subset)
            if self.screen is not None:
                res = self.screen(cand)
                cv = np.full(cand.shape[1], np.inf)
                if isinstance(res, tuple):          # the screen evaluated its survivors itself
                    idx = np.asarray(res[0], dtype=int)
                    self.evals += idx.size
                    cv[idx] = res[1]
                else:
                    idx = np.asarray(res, dtype=int)
                    if idx.size:
                        cv[idx] = self._eval(cand[:, idx])
            else:
                cv = self._eval(cand)
            better = cv < val
            pop[:, better] = cand[:, better]
            val[better] = cv[better]
            mean, spread_mean, spread_max = stats()
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
        return DEResult(gbest_val, gbest_vec, hist, self.evals, time.perf_counter() - t0, pop, val)
