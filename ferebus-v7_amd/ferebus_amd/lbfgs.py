"""L-BFGS polishing of one position on the cost -logL (mirror of
src/optimiser/fb_lbfgs.f90), driven by the library's adjoint gradient.

The reference PSO has no gradient step (gpr.f90:272-282 parses the flags but
never uses them); polishing is an optional addition, off by default.  Two-loop
recursion (Nocedal 1980), H0 = (s'y / y'y) I from the newest pair (1/|g| on
the first step), backtracking Armijo line search (c1 = 1e-4, halving); a
trial point whose matrix is not positive definite counts as a failed step.
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


def _cost_grad(gp, x: np.ndarray) -> Tuple[float, np.ndarray, bool]:
    ll, g, info = gp.swarm_grad(x[None, :])
    return -float(ll[0]), -g[0], bool(info[0] == 0)


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
        t, ok = 1.0, False
        for _ in range(cfg.max_backtracks + 1):
            xt = x + t * d
            ft, gt, pd = _cost_grad(gp, xt)
            evals += 1
            ok = pd and ft <= f + cfg.c1 * t * gd
            if ok:
                break
            t *= 0.5
        if not ok:
            break
        sy = (xt - x) @ (gt - g)
        if sy > 0.0:
            hist.append((xt - x, gt - g, 1.0 / sy))
        x, f, g = xt, ft, gt
        accepted += 1
    return x, f, evals, accepted
