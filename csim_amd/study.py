"""Surrogate-screened differential evolution against plain DE (no reference counterpart).

The hex-FPM optimisation of csim (differential evolution over the sag vector,
src/optimization/differentialEvolution/, objective = dark-hole mean contrast,
src/tools/regionContrast.cpp:145-205) run two ways on the batched GPU chain:

  plain      every mutant of a generation is propagated (9 wavelengths each) and
             selected on its exact dark-hole contrast;
  screened   a 512-mutant generation is ranked by the MLP surrogate (hex
             thicknesses + wavelength -> 32 annulus log10-contrasts) and only the
             top 32 are propagated; the exact values of those survivors decide the
             selection, are the surrogate's measured error, and are added to its
             training set (a few fine-tuning steps per generation).

The surrogate is trained first on propagations of random thickness vectors
(t ~ U(0, 1 um)) made by the same chain.  Usage:

    python -m csim_amd.study [--n 512] [--train 200704] [--epochs 10] [--seconds 20]

prints one JSON object (data-generation and training times, held-out rank
correlation, and the DE traces with the time each reached the plain run's
final contrast).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import workloads as W
from .chain import Chain
from .de import DEConfig, DifferentialEvolution
from .surrogate import MLP, N_BINS, annulus_bins, features_device, objective_from_profiles


class ExactObjective:
    """Dark-hole mean contrast of every candidate of a population, in fixed-size
    candidate batches (one compiled chain per batch size).  Optionally returns the
    per-(candidate, wavelength) annulus means used as surrogate labels."""

    def __init__(self, ctx, workload, batch, precision="bf16x3"):
        self.w = workload
        self.batch = batch
        self.ch = Chain(ctx, workload.script)
        W.apply(self.ch, workload)
        self.ch.set_precision(precision)
        self.nhex = workload.meta["n_hex"]
        self.ch.set_candidates("fpm", np.zeros((batch, self.nhex)))
        self.ch.contrast()                          # compile and calibrate once
        self.bins, self.counts = annulus_bins(self.ch)
        self.peak = self.ch.contrast()[1]["calib_peak"]
        self.nl = self.ch.info()["n_lambdas"]
        self.calls = 0

    def evaluate(self, thickness, with_profiles=False):
        """thickness: (P, n_hex) metres -> contrast (P,) [, labels (P, n_lambda, N_BINS)]."""
        P = thickness.shape[0]
        vals = np.empty(P)
        prof = np.empty((P, self.nl, N_BINS)) if with_profiles else None
        for i in range(0, P, self.batch):
            chunk = thickness[i:i + self.batch]
            k = chunk.shape[0]
            if k < self.batch:                      # pad with copies: the batch size stays fixed
                chunk = np.concatenate([chunk, np.repeat(chunk[-1:], self.batch - k, axis=0)])
            self.ch.set_candidates("fpm", W.thickness_to_sag(chunk))
            c, _ = self.ch.contrast()
            vals[i:i + k] = c[:k]
            if with_profiles:
                p = self.ch.field_profiles(self.bins, N_BINS).reshape(self.batch, self.nl, N_BINS)
                prof[i:i + k] = p[:k] / self.counts[None, None, :] / self.peak
            self.calls += k
        return (vals, prof) if with_profiles else vals

    def __call__(self, pop):            # DE interface: pop (D, P), numpy or a torch tensor
        if not isinstance(pop, np.ndarray):
            pop = pop.cpu().numpy()
        return self.evaluate(np.ascontiguousarray(pop.T))


def log_labels(prof):
    return np.log10(np.maximum(prof, 1e-14)).astype(np.float32)


def spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


class ScreenedSurrogate:
    """Ranks mutants by predicted objective, keeps the top `keep`, and learns from the
    exact values of the survivors: their labels go into a device ring buffer and a few
    Adam steps on random batches of it follow every generation.

    The exact propagation checks the surrogate: `n_control` of the `keep` exact evaluations of a
    generation go to mutants drawn at random from those the surrogate did not pick, and
    the generation's screening gain is the mean log10 contrast of those controls minus
    that of the picked ones (positive: the surrogate's picks are darker than chance).
    When its median over the last `fallback_window` generations falls below
    `fallback_gain` (picks clearly brighter than chance: the surrogate misleads the
    search) it is no longer trusted: `fallback` turns true, and the driver continues
    as plain DE (compare_de); until then a falling-back hook evaluates every mutant
    exactly.  Within a converged population picks and controls are alike, their gain
    is noise around zero, and the surrogate stays.  The rank correlation of predicted and exact
    contrast over the evaluated mutants is recorded too; within a converged population
    it is close to zero even for a useful surrogate, so it is reported, not acted on."""

    def __init__(self, mlp, lambdas, counts, exact, keep=32, tune_steps=2, lr=3e-4, capacity=32768, seed=5,
                 n_control=4, fallback_window=8, fallback_gain=-0.02, max_generations=None):
        import torch
        self.mlp, self.lambdas, self.counts, self.exact = mlp, list(lambdas), counts, exact
        self.keep, self.tune_steps, self.lr = keep, tune_steps, lr
        self.Xd = torch.empty((capacity, mlp.n_in), dtype=torch.float32, device="cuda")
        self.Yd = torch.empty((capacity, mlp.n_out), dtype=torch.float32, device="cuda")
        self.cap, self.pos, self.filled = capacity, 0, 0
        self.gen = torch.Generator(device="cuda")
        self.gen.manual_seed(seed)
        self.errors = []                # per generation: Spearman(pred, exact) of the survivors
        self.seconds = {"predict": 0.0, "exact": 0.0, "learn": 0.0}   # host wall time per phase
        self.n_control, self.fallback_window, self.fallback_gain = n_control, fallback_window, fallback_gain
        self.max_generations = max_generations   # screened generations before plain DE takes over (None: no limit)
        self.gains = []                 # per generation: screening gain (log10 contrast, controls - picks)
        self.rng = np.random.Generator(np.random.Philox(seed))
        self.generation = 0
        self.fallback_generation = None     # generation from which every mutant is evaluated exactly

    def predict(self, thickness):
        import torch
        if not hasattr(self, "_w"):
            self._w = torch.from_numpy(self.counts / self.counts.sum()).float().cuda()
        y = self.mlp.forward(features_device(self.mlp.ctx, thickness, self.lambdas))
        C = torch.pow(10.0, y).reshape(thickness.shape[0], len(self.lambdas), -1)
        return (C.sum(dim=1) @ self._w).double().cpu().numpy()

    def _push(self, X, Y):
        import torch
        n = X.shape[0]
        idx = (self.pos + np.arange(n)) % self.cap
        it = torch.from_numpy(idx).cuda()
        self.Xd.index_copy_(0, it, X)
        self.Yd.index_copy_(0, it, torch.from_numpy(Y).cuda())
        self.pos = (self.pos + n) % self.cap
        self.filled = min(self.cap, self.filled + n)

    def tune(self):
        """A few Adam steps on random batches of the ring buffer."""
        import torch
        nb = min(1024, self.filled)
        for _ in range(self.tune_steps):
            sel = torch.randint(0, self.filled, (nb,), device="cuda", generator=self.gen)
            self.mlp.train_step(self.Xd.index_select(0, sel), self.Yd.index_select(0, sel), self.lr, want_loss=False)

    @property
    def fallback(self):
        return self.fallback_generation is not None

    def screen_and_eval(self, pop):
        """DE hook: pop (D, P), numpy or a device tensor -> (indices evaluated, their
        exact values).  A device population stays there for the prediction and the
        features; only the evaluated candidates go to the host."""
        import torch
        self.generation += 1
        t0 = time.perf_counter()
        dev = torch.is_tensor(pop)
        t = pop.T.contiguous() if dev else np.ascontiguousarray(pop.T)
        host = (lambda x: x.cpu().numpy()) if dev else (lambda x: x)
        if self.fallback:                           # plain DE: every mutant exactly
            vals = self.exact.evaluate(host(t))
            self.seconds["exact"] += time.perf_counter() - t0
            return np.arange(t.shape[0]), vals
        pred = self.predict(t)
        order = np.argsort(pred)
        npick = self.keep - self.n_control
        ctrl = self.rng.choice(order[npick:], size=min(self.n_control, order.size - npick), replace=False)
        idx = np.concatenate([order[:npick], ctrl])
        sel = t[torch.as_tensor(idx, device=t.device)] if dev else t[idx]
        t1 = time.perf_counter()
        vals, prof = self.exact.evaluate(host(sel), with_profiles=True)
        t2 = time.perf_counter()
        if idx.size > 2:
            self.errors.append(spearman(pred[idx], vals))
        if ctrl.size:
            lv = np.log10(np.maximum(vals, 1e-300))
            self.gains.append(float(lv[npick:].mean() - lv[:npick].mean()))
            recent = self.gains[-self.fallback_window:]
            if len(recent) == self.fallback_window and np.median(recent) < self.fallback_gain:
                self.fallback_generation = self.generation + 1
        if self.max_generations is not None and self.generation >= self.max_generations and not self.fallback:
            self.fallback_generation = self.generation + 1
        self._push(features_device(self.mlp.ctx, sel, self.lambdas), log_labels(prof).reshape(-1, N_BINS))
        self.tune()
        t3 = time.perf_counter()
        self.seconds["predict"] += t1 - t0
        self.seconds["exact"] += t2 - t1
        self.seconds["learn"] += t3 - t2
        return idx, vals


def train_surrogate(ctx, w, n_train=22528, epochs=20, precision="bf16x3", seed=20261207, keep=32, log=print):
    """Propagates n_train random thickness vectors, trains the MLP on 90 % of them
    and measures it on the rest.  Returns (ScreenedSurrogate, report dict)."""
    import torch
    nhex, n_lambda = w.meta["n_hex"], len(w.meta["lambdas"])
    out = {}
    gen = ExactObjective(ctx, w, 256, precision)
    lambdas = gen.ch.lambdas()
    rng = np.random.Generator(np.random.Philox(seed))
    T = rng.random((n_train, nhex)) * 1e-6
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vals, prof = gen.evaluate(T, with_profiles=True)
    torch.cuda.synchronize()
    out["data_seconds"] = time.perf_counter() - t0
    out["data_propagations"] = int(n_train * n_lambda)
    log(f"data: {n_train} candidates x {n_lambda} lambda in {out['data_seconds']:.2f} s")
    # profiles reproduce the exact objective (consistency of the labels)
    rebuilt = objective_from_profiles(prof, gen.counts)
    out["label_consistency_max_rel"] = float(np.max(np.abs(rebuilt - vals) / vals))
    nh = n_train // 10
    X = features_device(ctx, T, lambdas)
    Y = torch.from_numpy(log_labels(prof).reshape(-1, N_BINS)).cuda()
    ntr = (n_train - nh) * n_lambda
    mlp = MLP(ctx, X.shape[1], 512, 512, N_BINS, seed=seed & 0xffff)
    t0 = time.perf_counter()
    hist = mlp.fit(X[:ntr], Y[:ntr], epochs=epochs, batch=1024, lr=1e-3, seed=seed)
    torch.cuda.synchronize()
    out["train_seconds"] = time.perf_counter() - t0
    out["train_loss"] = [float(hist[0]), float(hist[-1])]
    sur = ScreenedSurrogate(mlp, lambdas, gen.counts, None, keep=keep)
    pred = sur.predict(T[n_train - nh:])
    held = vals[n_train - nh:]
    out["holdout_spearman"] = spearman(pred, held)
    out["holdout_rel_err_median"] = float(np.median(np.abs(pred - held) / held))
    # screening gain on random candidates: mean exact contrast of the surrogate's
    # top 5 % against the mean of all
    top = np.argsort(pred)[:max(1, nh // 20)]
    out["holdout_top5pct_mean_over_mean"] = float(held[top].mean() / held.mean())
    log(f"train {out['train_seconds']:.2f} s, loss {hist[0]:.4f} -> {hist[-1]:.4f}, "
        f"held-out Spearman {out['holdout_spearman']:.3f}")
    gen.ch.close()
    return sur, out


def compare_de(ctx, w, sur, seconds=None, eval_budget=None, plain_pop=64, screen_pop=512, precision="bf16x3",
               seed=20261207, log=print, screen_generations=None):
    """Plain DE, then screened DE (stopping at the plain run's best value when a time
    budget is given) starting from the trained model, which is restored afterwards.
    When the exact check drops the surrogate (or after screen_generations screened
    generations) the screened run continues as plain DE from its best members.
    Budgets: wall seconds or exact candidate evaluations."""
    import torch
    nhex = w.meta["n_hex"]
    theta = sur.mlp.get_params()
    out = {}
    cfg = DEConfig(population_size=plain_pop, global_bounds=(0.0, 1e-6), init_bounds=(0.0, 1e-6),
                   max_iterations=10 ** 9, seed=seed + 1)
    exact_p = ExactObjective(ctx, w, plain_pop, precision)
    de = DifferentialEvolution(exact_p, nhex, cfg)
    torch.cuda.synchronize()
    rp = de.run(time_budget=seconds, eval_budget=eval_budget)
    out["plain"] = {"population": plain_pop, "best": rp.best_value, "seconds": rp.seconds,
                    "evaluations": rp.evaluations, "generations": len(rp.history) - 1}
    log(f"plain DE: best {rp.best_value:.4e} after {rp.seconds:.1f} s, {rp.evaluations} candidates")
    exact_p.ch.close()
    target = rp.best_value if seconds is not None else None
    sur.mlp.set_params(theta)
    exact_s = ExactObjective(ctx, w, sur.keep, precision)
    s1 = ScreenedSurrogate(sur.mlp, sur.lambdas, sur.counts, exact_s, keep=sur.keep,
                           max_generations=screen_generations)
    cfg_s = DEConfig(population_size=screen_pop, global_bounds=(0.0, 1e-6), init_bounds=(0.0, 1e-6),
                     max_iterations=10 ** 9, seed=seed + 2)
    # the screened population (512 mutants a generation) lives on the GPU
    de_s = DifferentialEvolution(exact_s, nhex, cfg_s, screen=s1.screen_and_eval, device="cuda")
    rs = de_s.run(time_budget=seconds, target=target, eval_budget=eval_budget, stop=lambda: s1.fallback)
    best_before_fallback = rs.best_value
    if s1.fallback:
        # continue as plain DE from the screened population's best members, for what is
        # left of the budgets; its history continues the screened run's clock and counts
        order = np.argsort(rs.values)[:plain_pop]
        exact_f = ExactObjective(ctx, w, plain_pop, precision)
        de_f = DifferentialEvolution(exact_f, nhex, DEConfig(population_size=plain_pop, global_bounds=(0.0, 1e-6),
                                                             init_bounds=(0.0, 1e-6), max_iterations=10 ** 9,
                                                             seed=seed + 3))
        left_s = None if seconds is None else max(0.0, seconds - rs.seconds)
        left_e = None if eval_budget is None else max(0, eval_budget - rs.evaluations)
        rf = de_f.run(time_budget=left_s, target=target, eval_budget=left_e, init_pop=rs.population[:, order],
                      init_val=rs.values[order])
        exact_f.ch.close()
        hist = rs.history + [(g + rs.history[-1][0], t + rs.seconds, min(v, rs.best_value), e + rs.evaluations)
                             for g, t, v, e in rf.history]
        rs = type(rs)(min(rs.best_value, rf.best_value),
                      rf.best_vector if rf.best_value < rs.best_value else rs.best_vector, hist,
                      rs.evaluations + rf.evaluations, rs.seconds + rf.seconds)
    reach = next((h for h in rs.history if h[2] <= rp.best_value), None)
    out["screened"] = {"population": screen_pop, "keep": s1.keep, "best": rs.best_value, "seconds": rs.seconds,
                       "evaluations": rs.evaluations, "generations": len(rs.history) - 1,
                       "seconds_to_plain_best": reach[1] if reach else None,
                       "evaluations_to_plain_best": reach[3] if reach else None,
                       "survivor_spearman_median": float(np.median(s1.errors)) if s1.errors else None,
                       "fallback_generation": s1.fallback_generation,
                       "best_before_fallback": best_before_fallback,
                       "screening_gain_median": float(np.median(s1.gains)) if s1.gains else None,
                       "seconds_by_phase": {k: round(v, 3) for k, v in s1.seconds.items()}}
    out["wall_ratio_vs_plain"] = (reach[1] / rp.seconds) if reach else None
    log(f"screened DE: best {rs.best_value:.4e} after {rs.seconds:.1f} s; reached plain best: "
        f"{(reach[1], reach[3]) if reach else None}")
    exact_s.ch.close()
    sur.mlp.set_params(theta)
    return out


def run(n=512, n_train=200704, seconds=20.0, n_lambda=9, plain_pop=64, screen_pop=512, keep=32, epochs=10,
        precision="bf16x3", seed=20261207, log=print):
    from .ops import Context
    ctx = Context(0)
    w = W.workload3(n=n, n_lambda=n_lambda)
    out = {"pupil": n, "n_lambda": n_lambda, "n_hex": w.meta["n_hex"], "precision": precision}
    sur, tr = train_surrogate(ctx, w, n_train, epochs, precision, seed, keep, log)
    out.update(tr)
    out.update(compare_de(ctx, w, sur, seconds=seconds, plain_pop=plain_pop, screen_pop=screen_pop,
                          precision=precision, seed=seed, log=log))
    if out.get("wall_ratio_vs_plain") is not None:
        out["wall_ratio_incl_training"] = (out["screened"]["seconds_to_plain_best"] + out["data_seconds"] +
                                           out["train_seconds"]) / out["plain"]["seconds"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--train", type=int, default=200704)
    ap.add_argument("--seconds", type=float, default=20.0)
    ap.add_argument("--lambdas", type=int, default=9)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--seed", type=int, default=20261207, help="training data, model and DE streams")
    a = ap.parse_args()
    res = run(n=a.n, n_train=a.train, seconds=a.seconds, n_lambda=a.lambdas, epochs=a.epochs, precision=a.precision,
              seed=a.seed, log=lambda s: print(s, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
