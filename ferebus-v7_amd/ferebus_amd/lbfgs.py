"""L-BFGS polishing of one position on the cost -logL (mirror of
src/optimiser/fb_lbfgs.f90), driven by the library's adjoint gradient.

The reference PSO has no gradient step (gpr.f90:272-282 parses the flags but
never uses them); polishing is an optional addition, off by default.  Two-loop
recursion (Nocedal 1980), H0 = (s'y / y'y) I from the newest pair (1/|g| on
the first step), backtracking Armijo line search (c1 = 1e-4, halving); a
trial point whose matrix is not positive definite counts as a failed step.
The full step t = 1 is tried alone; after a rejection the trial steps
1/2, 1/4, ... are evaluated `batch` at a time in one swarm gradient call (a
swarm of a few particles costs about as much as one on the GPU).  The
accepted step is the first of the sequence that satisfies the Armijo
condition, exactly as in a one-at-a-time backtracking search.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass
from typing import Tuple

import numpy as np


@dataclass
class LbfgsSettings:
    steps: int = 5
    memory: int = 5
    c1: float = 1e-4
    max_backtracks: int = 30
    batch: int = 4           # trial steps per gradient call


def _cost_grad(gp, x: np.ndarray) -> Tuple[float, np.ndarray, bool]:
    ll, g, info = gp.swarm_grad(x[None, :])
    return -float(ll[0]), -g[0], bool(info[0] == 0)


def _cost_grad_batch(gp, xs: np.ndarray):
    ll, g, info = gp.swarm_grad(xs)
    return -ll, -g, info == 0


def polish(gp, x0, cfg: LbfgsSettings = LbfgsSettings()):
    """(x, cost, evals, accepted): cost = -logL at the polished x."""
    x = np.array(x0, dtype=np.float64)
    f, g, ok = _cost_grad(gp, x)
    evals, accepted = 1, 0
    if not ok:
        return x, np.inf, evals, accepted
    hist: deque = deque(maxlen=max(cfg.memory, 1))
    while accepted < cfg.steps:
        q = g.copy()
        alphas = []
        for s, y, rho in reversed(hist):
            a = rho * (s @ q)
            alphas.append(a)
            q -= a * y
        if hist:
            s, y, _ = hist[-1]
            gamma = (s @ y) / (y @ y)
        else:
            gamma = 1.0 / max(np.linalg.norm(g), np.finfo(float).tiny)
        q *= gamma
        for (s, y, rho), a in zip(hist, reversed(alphas)):
            b = rho * (y @ q)
            q += (a - b) * s
        d = -q
        gd = g @ d
        if not gd < 0.0:
            d = -gamma * g
            gd = g @ d
            if not gd < 0.0:
                break
        ok, k0 = False, 0
        nb = max(1, min(cfg.batch, getattr(gp, "max_swarm", 1)))
        while not ok and k0 <= cfg.max_backtracks:
            # the full step alone first (usually accepted), then `nb` at a time
            ts = 0.5 ** np.arange(k0, min(k0 + (1 if k0 == 0 else nb), cfg.max_backtracks + 1))
            fs, gs, pds = _cost_grad_batch(gp, x[None, :] + ts[:, None] * d[None, :])
            evals += len(ts)
            for j, t in enumerate(ts):
                if pds[j] and fs[j] <= f + cfg.c1 * t * gd:
                    ok = True
                    xt, ft, gt = x + t * d, float(fs[j]), gs[j].copy()
                    break
            k0 += len(ts)
        if not ok:
            break
        sy = (xt - x) @ (gt - g)
        if sy > 0.0:
            hist.append((xt - x, gt - g, 1.0 / sy))
        x, f, g = xt, ft, gt
        accepted += 1
    return x, f, evals, accepted
