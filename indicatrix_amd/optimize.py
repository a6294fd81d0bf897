"""Surrogate-screened optimisation of a cut for a given material.

The reference's Optimize (crates/indicatrix-cut-core/src/optimize/mod.rs:8-35)
scores every candidate with the metric tracer (evaluate_objective,
crates/indicatrix-cut-core/src/optimize/objective.rs:303).  This loop keeps
the tracer as the judge but screens each step's candidates with the BF16
surrogate first:

  1. propose n_candidates seeded cuts around the best verified cut
     (uniform over the family box on the first step);
  2. score all of them x every tilt with the surrogate in one batched pass;
  3. trace the top n_verify with the metric tracer (every tilt) and keep the
     best verified one;
  4. optionally refine the surrogate on every design verified so far
     (refine_steps Adam steps), so that its error is lowest where the search
     is heading; the error reported for a step is always measured before the
     step's own designs are added.

The objective is the equal-weight mean of the three tilt-averaged metrics
(brilliance %, fire index, scintillation %), following the reference's
equal default weights (objective.rs:131); higher is better.  The
tracer-only loop (TracerOnlyOptimizer) is the same proposal rule with every
candidate traced, the method without the surrogate.
"""
import time

import numpy as np

from .designs import TILTS_C3
from .surrogate import design_features, predict_designs, relative_error, trace_metrics

VERIFY_SHAPE = (128, 128, 64)   # width, height, samples of a verification render


def objective(metrics):
    """(..., tilts, 3) metrics -> (...) scores."""
    return np.asarray(metrics, np.float64).mean(axis=-2).mean(axis=-1)


class _Loop:
    def __init__(self, backend, nd, abbe, seed, n_candidates, tilts, verify_shape, sigma0, sigma_decay):
        self.backend = backend
        self.nd = float(nd)
        self.abbe = float(abbe)
        self.rng = np.random.default_rng(seed)
        self.n_candidates = int(n_candidates)
        self.tilts = np.asarray(tilts, np.float32)
        self.verify_shape = verify_shape
        self.sigma = float(sigma0)
        self.sigma_decay = float(sigma_decay)
        self.best_u = None
        self.best_score = -np.inf
        self.history = []

    def _propose(self):
        n = self.n_candidates
        if self.best_u is None:
            u = self.rng.random((n, 13))
        else:
            u = np.clip(self.best_u[None, :] + self.sigma * self.rng.standard_normal((n, 13)), 0.0, 1.0)
            u[0] = self.best_u          # the incumbent stays in the pool
            self.sigma *= self.sigma_decay
        return u

    def _trace(self, u):
        w, h, s = self.verify_shape
        n = len(u)
        return trace_metrics(self.backend, u, np.full(n, self.nd), np.full(n, self.abbe), self.tilts,
                             w, h, s)

    def _accept(self, u, scores):
        i = int(np.argmax(scores))
        if scores[i] > self.best_score:
            self.best_score = float(scores[i])
            self.best_u = u[i].copy()


class ScreenedOptimizer(_Loop):
    """Surrogate screening + tracer verification of the top n_verify."""

    def __init__(self, backend, net, nd, abbe, seed=0, n_candidates=8192, n_verify=64, tilts=TILTS_C3,
                 verify_shape=VERIFY_SHAPE, sigma0=0.15, sigma_decay=0.93, refine_steps=100,
                 refine_lr=1e-4):
        super().__init__(backend, nd, abbe, seed, n_candidates, tilts, verify_shape, sigma0, sigma_decay)
        self.net = net
        self.n_verify = int(n_verify)
        self.refine_steps = int(refine_steps)
        self.refine_lr = float(refine_lr)
        self._seen_x = []
        self._seen_t = []

    def step(self):
        t0 = time.perf_counter()
        u = self._propose()
        n = len(u)
        pred = predict_designs(self.net, u, np.full(n, self.nd), np.full(n, self.abbe), self.tilts)
        t1 = time.perf_counter()
        top = np.argsort(-objective(pred), kind="stable")[:self.n_verify]
        traced = self._trace(u[top])
        t2 = time.perf_counter()
        scores = objective(traced)
        self._accept(u[top], scores)
        if self.refine_steps:
            k = len(top)
            self._seen_x.append(design_features(u[top], np.full(k, self.nd), np.full(k, self.abbe), self.tilts))
            self._seen_t.append(traced.reshape(-1, traced.shape[-1]))
            self.net.fine_tune(np.concatenate(self._seen_x), np.concatenate(self._seen_t),
                               self.refine_steps, self.refine_lr, seed=len(self.history))
        t3 = time.perf_counter()
        rec = dict(screen_s=t1 - t0, verify_s=t2 - t1, refine_s=t3 - t2, step_s=t3 - t0, candidates=n,
                   verified=len(top), best=self.best_score, step_best=float(scores.max()),
                   rel_err=relative_error(pred[top], traced),
                   obj_rel_err=np.abs(objective(pred[top]) - scores) / np.abs(scores),
                   pred=pred[top], traced=traced)
        self.history.append(rec)
        return rec


class TracerOnlyOptimizer(_Loop):
    """The same proposals with every candidate traced (no surrogate)."""

    def __init__(self, backend, nd, abbe, seed=0, n_candidates=8192, tilts=TILTS_C3,
                 verify_shape=VERIFY_SHAPE, sigma0=0.15, sigma_decay=0.93):
        super().__init__(backend, nd, abbe, seed, n_candidates, tilts, verify_shape, sigma0, sigma_decay)

    def step(self):
        t0 = time.perf_counter()
        u = self._propose()
        traced = self._trace(u)
        t1 = time.perf_counter()
        scores = objective(traced)
        self._accept(u, scores)
        rec = dict(step_s=t1 - t0, candidates=len(u), verified=len(u), best=self.best_score,
                   step_best=float(scores.max()), traced=traced)
        self.history.append(rec)
        return rec
