"""Particle swarm optimisation of the GP hyperparameters on the device swarm
(the fb_pso_* entry points), mirror of src/optimiser/fb_pso.f90.

Semantics of the reference PSO (src/optimiser/pso/pso.f90,
src/optimiser/pso/particle.f90, src/optimiser/iter_state.f90):
  * initial swarm: positions U[lo, hi], velocities U[lo - hi, hi - lo]
    (pso.f90:291-292), evaluated once without retries (pso.f90:294-308);
  * each iteration: r1, r2 ~ U[0,1)^(S x P) drawn before the swarm loop
    (pso.f90:405-409); the velocity / position update, the evaluation and
    the selection of the previous iteration run on the device as one graph
    replay (fb_pso_step);
  * a particle whose matrix is not positive definite is re-randomised and
    re-evaluated until it succeeds (get_cost, particle.f90:149-162) -- all
    failed particles of an iteration together (fb_pso_retry);
  * pbest with <=, gbest in particle order with <= (pso.f90:440-459);
  * stopping: relative change of the best cost at most `tolerance` for
    `stall_iterations` iterations (pso.f90:619-636), or the iteration cap.
Optional (lbfgs_every > 0): every lbfgs_every iterations the global best is
polished by a few L-BFGS steps on the adjoint gradient and replaces gbest
when its cost is lower.  The RNG is numpy's PCG64 (the reference's Fortran
RNG stream cannot be matched; parity is checked by replaying the traced
positions).
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import _abi
from .lbfgs import LbfgsSettings, polish


@dataclass
class PsoSettings:
    swarm_size: int = 50
    iterations: int = 1000
    inertia_weight: float = 0.789              # config/optimiser/pso/pso.f90:48-53
    cognitive_learning_rate: float = 1.494
    social_learning_rate: float = 1.494
    relative_change: bool = False
    tolerance: float = 1e-6                    # stopping_criteria.f90:60-62
    stall_iterations: int = 20
    seed: int = 1
    max_retries: int = 10000
    search_min: float = 0.0                    # optimiser.f90(config):35-38
    search_max: float = 1.0
    noise_max: float = 1e-6                    # optimised nugget's upper bound (gpr.f90:229, 242 hard-code 1e-6)
    lbfgs_every: int = 0
    lbfgs: LbfgsSettings = field(default_factory=LbfgsSettings)


@dataclass
class PsoResult:
    best_position: np.ndarray
    best_cost: float
    iterations: int = 0
    evaluations: int = 0
    rejections: int = 0
    best_iter: int = 0
    seconds: float = 0.0
    polish_calls: int = 0
    polish_evals: int = 0
    polish_improvements: int = 0
    # (iteration, seconds since the start of the optimisation, gbest cost)
    history: List[Tuple[int, float, float]] = field(default_factory=list)


class DeviceSwarm:
    """The fb_pso swarm state of one SwarmLikelihood (positions, velocities,
    personal / global bests in device memory)."""

    def __init__(self, gp, s: int, w: float, c1: float, c2: float):
        self._lib = _abi.load()
        self.gp = gp
        self.S, self.P = int(s), gp.nparams
        self._p = ctypes.c_void_p()
        _abi.check(self._lib.fb_pso_create(ctypes.byref(self._p), gp._h, self.S, w, c1, c2), "fb_pso_create")

    def close(self) -> None:
        if self._p:
            self._lib.fb_pso_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rows(self, a, k) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (k, self.P):
            raise ValueError(f"expected shape ({k}, {self.P}), got {a.shape}")
        return a

    def init(self, x0, v0) -> Tuple[np.ndarray, np.ndarray]:
        x0, v0 = self._rows(x0, self.S), self._rows(v0, self.S)
        cost, info = np.empty(self.S), np.empty(self.S, dtype=np.int32)
        _abi.check(self._lib.fb_pso_init(self._p, _abi.dptr(x0), _abi.dptr(v0), _abi.dptr(cost),
                                         _abi.iptr(info)), "fb_pso_init")
        return cost, info

    def step(self, r1, r2) -> Tuple[np.ndarray, np.ndarray]:
        r1, r2 = self._rows(r1, self.S), self._rows(r2, self.S)
        cost, info = np.empty(self.S), np.empty(self.S, dtype=np.int32)
        _abi.check(self._lib.fb_pso_step(self._p, _abi.dptr(r1), _abi.dptr(r2), _abi.dptr(cost),
                                         _abi.iptr(info)), "fb_pso_step")
        return cost, info

    def retry(self, idx, x, v) -> Tuple[np.ndarray, np.ndarray]:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        k = len(idx)
        x, v = self._rows(x, k), self._rows(v, k)
        cost, info = np.empty(k), np.empty(k, dtype=np.int32)
        _abi.check(self._lib.fb_pso_retry(self._p, k, _abi.iptr(idx), _abi.dptr(x), _abi.dptr(v),
                                          _abi.dptr(cost), _abi.iptr(info)), "fb_pso_retry")
        return cost, info

    def select(self) -> Tuple[float, np.ndarray, int]:
        gcost = ctypes.c_double()
        gbest = np.empty(self.P)
        gidx = ctypes.c_int()
        _abi.check(self._lib.fb_pso_select(self._p, ctypes.byref(gcost), _abi.dptr(gbest), ctypes.byref(gidx)),
                   "fb_pso_select")
        return gcost.value, gbest, gidx.value

    def set_gbest(self, gbest, gcost: float) -> None:
        g = np.ascontiguousarray(gbest, dtype=np.float64)
        _abi.check(self._lib.fb_pso_set_gbest(self._p, _abi.dptr(g), float(gcost)), "fb_pso_set_gbest")

    def positions(self) -> Tuple[np.ndarray, np.ndarray]:
        x, v = np.empty((self.S, self.P)), np.empty((self.S, self.P))
        _abi.check(self._lib.fb_pso_positions(self._p, _abi.dptr(x), _abi.dptr(v)), "fb_pso_positions")
        return x, v


def search_bounds(gp, cfg: PsoSettings) -> Tuple[np.ndarray, np.ndarray]:
    """search_min_gpr / search_max_gpr (gpr.f90:204-246): the kernel
    parameters in [search_min, search_max], an optimised nugget in
    [1e-12, noise_max] (the reference fixes noise_max = 1e-6, gpr.f90:229,
    242)."""
    lo = np.full(gp.nparams, cfg.search_min)
    hi = np.full(gp.nparams, cfg.search_max)
    if gp.optimise_noise:
        lo[-1], hi[-1] = 1e-12, cfg.noise_max
    return lo, hi


def run_pso(gp, cfg: PsoSettings, trace: Optional[Callable] = None,
            target_cost: Optional[float] = None) -> PsoResult:
    """Runs the PSO on `gp` (a SwarmLikelihood with max_swarm >= swarm_size).

    trace(it, round, k, info, cost, theta) is called for every evaluated
    particle (positions are read back only when tracing).  With
    target_cost the run also stops as soon as gbest cost <= target_cost.
    """
    S, P = cfg.swarm_size, gp.nparams
    rng = np.random.default_rng(cfg.seed)
    lo, hi = search_bounds(gp, cfg)
    sw = DeviceSwarm(gp, S, cfg.inertia_weight, cfg.cognitive_learning_rate, cfg.social_learning_rate)
    res = PsoResult(best_position=np.empty(P), best_cost=math.inf)
    try:
        t0 = time.perf_counter()
        x = lo + rng.random((S, P)) * (hi - lo)
        v = (lo - hi) + rng.random((S, P)) * (2.0 * (hi - lo))
        cost, info = sw.init(x, v)
        res.evaluations += S
        res.rejections += int(np.sum(info != 0))
        if trace:
            for k in range(S):
                trace(0, 0, k, int(info[k]), float(cost[k]), x[k])
        gcost = math.inf
        for c in cost:                        # state%update in particle order
            if c <= gcost:
                gcost = float(c)
        res.history.append((0, time.perf_counter() - t0, gcost))
        stall_cost, stall = math.inf, 0
        it = 1
        while True:
            r1 = rng.random((S, P))
            r2 = rng.random((S, P))
            cost, info = sw.step(r1, r2)
            res.evaluations += S
            if trace:
                xs, _ = sw.positions()
                for k in range(S):
                    trace(it, 0, k, int(info[k]), float(cost[k]), xs[k])
            todo = np.flatnonzero(info != 0)
            rnd = 0
            while len(todo):
                res.rejections += len(todo)
                rnd += 1
                if rnd > cfg.max_retries:
                    raise RuntimeError("PSO: particles stayed non-positive-definite after max_retries "
                                       "re-randomisations")
                # particle.f90:117-128
                xr = lo + rng.random((len(todo), P)) * (hi - lo)
                vr = (lo - hi) + rng.random((len(todo), P)) * (2.0 * (hi - lo))
                c, inf = sw.retry(todo, xr, vr)
                res.evaluations += len(todo)
                cost[todo] = c
                if trace:
                    for r, k in enumerate(todo):
                        trace(it, rnd, int(k), int(inf[r]), float(c[r]), xr[r])
                todo = todo[inf != 0]
            for c in cost:
                if c <= gcost:
                    gcost = float(c)
                    res.best_iter = it
            if cfg.lbfgs_every > 0 and it % cfg.lbfgs_every == 0:
                gdev, gbest, _ = sw.select()
                xp, fp, ne, na = polish(gp, gbest, cfg.lbfgs)
                res.polish_calls += 1
                res.polish_evals += ne
                if fp < gdev:
                    sw.set_gbest(xp, fp)
                    gcost = float(fp)
                    res.best_iter = it
                    res.polish_improvements += 1
                if trace:
                    trace(it, -1, -1, 0, float(fp), xp)
            res.iterations = it
            res.history.append((it, time.perf_counter() - t0, gcost))
            if cfg.relative_change:
                if abs((gcost - stall_cost) / stall_cost) <= cfg.tolerance:
                    stall += 1
                else:
                    stall, stall_cost = 0, gcost
                if stall >= cfg.stall_iterations:
                    break
            if target_cost is not None and gcost <= target_cost:
                break
            if it >= cfg.iterations:
                break
            it += 1
        gdev, gbest, _ = sw.select()
        res.seconds = time.perf_counter() - t0
        res.best_cost, res.best_position = gdev, gbest
        return res
    finally:
        sw.close()
