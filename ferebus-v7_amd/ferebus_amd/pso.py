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
    `stall_iterations` iterations (pso.f90:619-636), or the iteration cap;
  * swarm updaters (config/optimiser/pso/pso.f90:141-157): "global" (the
    above) or "forced" (pso.f90:485-558): particles updated one after
    another with their own r1, r2, each evaluated and selected before the
    next moves (gbest only when pbest improved), velocity components whose
    partial potential |v| + |gbest - x| falls below `delta` re-drawn in
    [-delta, delta]; forced runs on the host, one particle per evaluation;
  * parameter updaters (config/optimiser/pso/pso.f90:115-139,
    pso.f90:638-825): constant; unified / individual adaptive (learning
    automata over w, c1, c2, learning_automata.f90:48-116, rewarded when
    n_improved / n >= success_threshold in integer division, pso.f90:760);
    random unified / individual (w ~ U[0,1], c1 ~ U[0, b], c2 ~ U[0, b - c1],
    b = 24 (1 - w^2) / (7 - 5 w)); unified variants update after every swarm
    update, individual ones after every particle of the forced updater only
    (the global updater never calls them, pso.f90:445-447).
Several training sets (run_pso_multi): their swarms advance in lock step,
one fb_multi_pso_step per iteration evaluating all of them together, each
with its own RNG and host rules, so each gets the run run_pso gives it.
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
from typing import Callable, Dict, List, Optional, Sequence, Tuple

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
    swarm_updater: str = "global"                # config/optimiser/pso/pso.f90:141-157
    delta: float = 1e-7                          # config/optimiser/pso/swarm_updater.f90:58
    parameter_updater: str = "constant"          # config/optimiser/pso/pso.f90:115-139
    adaptive: "AdaptiveSettings" = None

    def __post_init__(self):
        if self.adaptive is None:
            self.adaptive = AdaptiveSettings()
        self.swarm_updater = swarm_updater_name(self.swarm_updater)
        self.parameter_updater = parameter_updater_name(self.parameter_updater)


@dataclass
class AdaptiveSettings:
    """AdaptiveParameterUpdaterConfig (config/optimiser/pso/parameter_updater.f90:62-75)."""
    min_inertia_weight: float = 0.0
    max_inertia_weight: float = 1.0
    n_inertia_weight: int = 10
    min_cognitive_learning_rate: float = 0.0
    max_cognitive_learning_rate: float = 4.0
    n_cognitive_learning_rate: int = 10
    min_social_learning_rate: float = 0.0
    max_social_learning_rate: float = 4.0
    n_social_learning_rate: int = 10
    reward_step_length: float = 0.01
    penalty_step_length: float = 0.01
    success_threshold: float = 0.5


def sanitise(name: str) -> str:
    """Spaces to underscores, lower case (utils/str.f90:85-98)."""
    return name.replace(" ", "_").lower()


_SWARM_UPDATERS = {"global": "global", "default": "global", "forced": "forced"}
_PARAMETER_UPDATERS = {"constant": "constant", "none": "constant",
                       "unified_adaptive": "unified_adaptive", "uapso": "unified_adaptive", "ua": "unified_adaptive",
                       "individual_adaptive": "individual_adaptive", "iapso": "individual_adaptive",
                       "ia": "individual_adaptive",
                       "random_unified": "random_unified", "random": "random_unified", "ru": "random_unified",
                       "random_individual": "random_individual", "ri": "random_individual"}
_STOPPING = {"none": False, "relative_change": True}


def swarm_updater_name(name: str) -> str:
    k = sanitise(name)
    if k not in _SWARM_UPDATERS:
        raise ValueError(f"unknown swarm_updater: {name!r} (global, default, forced)")
    return _SWARM_UPDATERS[k]


def parameter_updater_name(name: str) -> str:
    k = sanitise(name)
    if k not in _PARAMETER_UPDATERS:
        raise ValueError(f"unknown parameter_updater: {name!r}")
    return _PARAMETER_UPDATERS[k]


def stopping_is_relative_change(name: str) -> bool:
    """stopping_criteria (config/optimiser/pso/pso.f90:92-113): "none" or
    "relative_change"; "potential_based" and "sign_test" parse in the
    reference but have no implementation there (pso.f90:360-366), so they are
    refused here rather than silently run to the iteration cap."""
    k = sanitise(name)
    if k in ("potential_based", "sign_test"):
        raise ValueError(f"stopping_criteria {name!r} has no implementation in the reference either")
    if k not in _STOPPING:
        raise ValueError(f"unknown stopping_criteria: {name!r}")
    return _STOPPING[k]


class LearningAutomata:
    """learning_automata.f90:48-116: n actions spread evenly over [lo, hi],
    probabilities 1/n; choose() picks the last action whose closed interval
    [sum of the earlier probabilities, that sum + its own] holds
    u * (sum of all probabilities) (the loop's `continue` does not leave it);
    reward / penalise move the probabilities by the step lengths."""

    def __init__(self, lo, hi, n, reward_step, penalty_step):
        self.values = [lo + float(i) * (hi - lo) / float(n - 1) for i in range(n)]
        self.prob = [1.0 / float(n)] * n
        self.selected = 0
        self.reward_step, self.penalty_step = reward_step, penalty_step

    def choose(self, u: float) -> float:
        total = 0.0
        for q in self.prob:
            total += q
        rand = u * total
        prev = 0.0
        for i, q in enumerate(self.prob):
            cur = prev + q
            if prev <= rand <= cur:
                self.selected = i
            prev = prev + q
        return self.values[self.selected]

    def reward(self):
        a = self.reward_step
        self.prob = [q + a * (1.0 - q) if i == self.selected else q * (1.0 - a) for i, q in enumerate(self.prob)]

    def penalise(self):
        b, n = self.penalty_step, len(self.prob)
        self.prob = [q * (1.0 - b) if i == self.selected else b / float(n - 1) + (1.0 - b) * q
                     for i, q in enumerate(self.prob)]


class ParameterUpdater:
    """The swarm's (w, c1, c2) and how they change (pso.f90:638-825).
    `draws` records every uniform it consumed, in order."""

    def __init__(self, kind: str, cfg: "PsoSettings", rng: np.random.Generator, costs):
        self.kind = kind
        self.rng = rng
        self.coef = [cfg.inertia_weight, cfg.cognitive_learning_rate, cfg.social_learning_rate]
        self.per_swarm = kind in ("unified_adaptive", "random_unified")
        self.per_particle = kind in ("individual_adaptive", "random_individual")
        self.draws: List[float] = []
        if kind in ("unified_adaptive", "individual_adaptive"):
            a = cfg.adaptive
            self.automata = [
                LearningAutomata(a.min_inertia_weight, a.max_inertia_weight, a.n_inertia_weight,
                                 a.reward_step_length, a.penalty_step_length),
                LearningAutomata(a.min_cognitive_learning_rate, a.max_cognitive_learning_rate,
                                 a.n_cognitive_learning_rate, a.reward_step_length, a.penalty_step_length),
                LearningAutomata(a.min_social_learning_rate, a.max_social_learning_rate,
                                 a.n_social_learning_rate, a.reward_step_length, a.penalty_step_length)]
            self.threshold = a.success_threshold
            self.previous = [float(c) for c in costs]
            self._choose()
        elif kind in ("random_unified", "random_individual"):
            self._random()

    def _u(self) -> float:
        u = float(self.rng.random())
        self.draws.append(u)
        return u

    def _choose(self):
        self.coef = [la.choose(self._u()) for la in self.automata]

    def _random(self):
        # pso.f90:812-824 (inequality of doi:10.1007/s11721-017-0150-9)
        w = self._u()
        b = 24 * (1 - w * w) / (7 - 5 * w)
        c1 = 0.0 + self._u() * (b - 0.0)
        c2 = 0.0 + self._u() * ((b - c1) - 0.0)
        self.coef = [w, c1, c2]

    def update(self, costs, first: int = 0):
        """update_parameters over particles first .. first + len(costs) - 1."""
        if self.kind == "constant":
            return
        if self.kind in ("random_unified", "random_individual"):
            self._random()
            return
        n = len(costs)
        improved = 0
        for i, c in enumerate(costs):
            if c < self.previous[first + i]:       # pso.f90:752-757
                improved += 1
            self.previous[first + i] = float(c)
        if (improved // n) >= self.threshold:      # integer division, pso.f90:760
            for la in self.automata:
                la.reward()
        else:
            for la in self.automata:
                la.penalise()
        self._choose()


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


def _record(record, **kw):
    if record is not None:
        record.append({k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in kw.items()})


class _SwarmRun:
    """The host side of one PSO run around its device swarm (ParticleSwarm,
    pso.f90:382-472): the initial swarm, the draws of each iteration, the
    re-randomisation of failed particles (particle.f90:117-128), the
    running gbest, the parameter updater, polishing and the stopping rules.
    The iteration's device step itself is left to the caller -- the swarm's
    own fb_pso_step (run_pso) or its part of fb_multi_pso_step
    (run_pso_multi) -- so both drivers apply the same rules in the same
    order and give the same results."""

    def __init__(self, gp, cfg: PsoSettings, trace, target_cost, record):
        self.gp, self.cfg = gp, cfg
        self.trace, self.target_cost, self.record = trace, target_cost, record
        self.S, self.P = cfg.swarm_size, gp.nparams
        self.rng = np.random.default_rng(cfg.seed)
        self.lo, self.hi = search_bounds(gp, cfg)
        self.sw = DeviceSwarm(gp, self.S, cfg.inertia_weight, cfg.cognitive_learning_rate,
                              cfg.social_learning_rate)
        self.res = PsoResult(best_position=np.empty(self.P), best_cost=math.inf)
        self.done = False

    def start(self) -> None:
        S, P, rng, lo, hi, res = self.S, self.P, self.rng, self.lo, self.hi, self.res
        self.t0 = time.perf_counter()
        x = lo + rng.random((S, P)) * (hi - lo)
        v = (lo - hi) + rng.random((S, P)) * (2.0 * (hi - lo))
        cost, info = self.sw.init(x, v)
        res.evaluations += S
        res.rejections += int(np.sum(info != 0))
        if self.trace:
            for k in range(S):
                self.trace(0, 0, k, int(info[k]), float(cost[k]), x[k])
        _record(self.record, it=0, kind="init", x=x, v=v, cost=cost, info=info)
        self.gcost = math.inf
        for c in cost:                        # state%update in particle order
            if c <= self.gcost:
                self.gcost = float(c)
        res.history.append((0, time.perf_counter() - self.t0, self.gcost))
        self.upd = ParameterUpdater(self.cfg.parameter_updater, self.cfg, rng, cost)
        _record(self.record, it=0, kind="coef", coef=list(self.upd.coef), draws=list(self.upd.draws))
        self.stall_cost, self.stall = math.inf, 0
        self.it = 1

    def draw(self) -> Tuple[np.ndarray, np.ndarray]:
        """The coefficients and uniforms of the next iteration's update."""
        _abi.check(self.sw._lib.fb_pso_set_coefficients(self.sw._p, *self.upd.coef), "fb_pso_set_coefficients")
        self.r1 = self.rng.random((self.S, self.P))
        self.r2 = self.rng.random((self.S, self.P))
        return self.r1, self.r2

    def after_step(self, cost, info) -> None:
        """Everything after the iteration's evaluation; sets `done`."""
        cfg, res, sw, it, S, P = self.cfg, self.res, self.sw, self.it, self.S, self.P
        res.evaluations += S
        if self.trace or self.record is not None:
            xs, vs = sw.positions()
            if self.trace:
                for k in range(S):
                    self.trace(it, 0, k, int(info[k]), float(cost[k]), xs[k])
            _record(self.record, it=it, kind="step", coef=list(self.upd.coef), r1=self.r1, r2=self.r2, x=xs,
                    v=vs, cost=cost, info=info)
        todo = np.flatnonzero(info != 0)
        rnd = 0
        while len(todo):
            res.rejections += len(todo)
            rnd += 1
            if rnd > cfg.max_retries:
                raise RuntimeError("PSO: particles stayed non-positive-definite after max_retries "
                                   "re-randomisations")
            # particle.f90:117-128
            xr = self.lo + self.rng.random((len(todo), P)) * (self.hi - self.lo)
            vr = (self.lo - self.hi) + self.rng.random((len(todo), P)) * (2.0 * (self.hi - self.lo))
            c, inf = sw.retry(todo, xr, vr)
            res.evaluations += len(todo)
            cost[todo] = c
            if self.trace:
                for r, k in enumerate(todo):
                    self.trace(it, rnd, int(k), int(inf[r]), float(c[r]), xr[r])
            _record(self.record, it=it, kind="retry", idx=todo.copy(), x=xr, v=vr, cost=c, info=inf)
            todo = todo[inf != 0]
        for c in cost:
            if c <= self.gcost:
                self.gcost = float(c)
                res.best_iter = it
        if self.upd.per_swarm:                # next_iter_pso, pso.f90:563-565
            self.upd.draws = []
            self.upd.update(cost)
            _record(self.record, it=it, kind="coef", coef=list(self.upd.coef), draws=list(self.upd.draws))
        if cfg.lbfgs_every > 0 and it % cfg.lbfgs_every == 0:
            gdev, gbest, _ = sw.select()
            xp, fp, ne, na = polish(self.gp, gbest, cfg.lbfgs)
            res.polish_calls += 1
            res.polish_evals += ne
            if fp < gdev:
                sw.set_gbest(xp, fp)
                self.gcost = float(fp)
                res.best_iter = it
                res.polish_improvements += 1
            if self.trace:
                self.trace(it, -1, -1, 0, float(fp), xp)
        res.iterations = it
        res.history.append((it, time.perf_counter() - self.t0, self.gcost))
        if cfg.relative_change:
            if abs((self.gcost - self.stall_cost) / self.stall_cost) <= cfg.tolerance:
                self.stall += 1
            else:
                self.stall, self.stall_cost = 0, self.gcost
            if self.stall >= cfg.stall_iterations:
                self.done = True
        if self.target_cost is not None and self.gcost <= self.target_cost:
            self.done = True
        if it >= cfg.iterations:
            self.done = True
        if not self.done:
            self.it += 1

    def finish(self) -> PsoResult:
        gdev, gbest, _ = self.sw.select()
        self.res.seconds = time.perf_counter() - self.t0
        self.res.best_cost, self.res.best_position = gdev, gbest
        return self.res

    def close(self) -> None:
        self.sw.close()


def run_pso(gp, cfg: PsoSettings, trace: Optional[Callable] = None,
            target_cost: Optional[float] = None, record: Optional[List[Dict]] = None) -> PsoResult:
    """Runs the PSO on `gp` (a SwarmLikelihood with max_swarm >= swarm_size).

    trace(it, round, k, info, cost, theta) is called for every evaluated
    particle (positions are read back only when tracing).  With
    target_cost the run also stops as soon as gbest cost <= target_cost.
    record (a list) receives one entry per event with everything the host
    drew and read back (coefficients, r1 / r2, re-randomised particles, the
    updater's uniforms, positions, costs, INFO), for replaying the rules.
    """
    if cfg.swarm_updater == "forced":
        return _run_forced(gp, cfg, trace, target_cost, record)
    run = _SwarmRun(gp, cfg, trace, target_cost, record)
    try:
        run.start()
        while not run.done:
            r1, r2 = run.draw()
            cost, info = run.sw.step(r1, r2)
            run.after_step(cost, info)
        return run.finish()
    finally:
        run.close()


def run_pso_multi(gps: Sequence, cfgs: Sequence[PsoSettings], traces: Optional[Sequence] = None,
                  target_costs: Optional[Sequence] = None,
                  records: Optional[Sequence] = None) -> List[PsoResult]:
    """The PSO of several training sets (one SwarmLikelihood each, any N) in
    lock step: every iteration advances all unfinished swarms with one
    fb_multi_pso_step call -- one captured graph that updates every swarm
    and evaluates all of them together (the schedules of
    fb_multi_swarm_loglik, the merged chained launches for uneven N) -- and
    then applies each run's host rules (retries, updater, polishing,
    stopping) as run_pso does.  A run that stops leaves the others going.
    Each result equals run_pso(gps[g], cfgs[g]) with the same settings, bit
    for bit.  Runs with the forced swarm updater (sequential by definition)
    run on their own first.  The reference trains its GPs one after another
    (app/main.f90:78-130)."""
    from .gpr import MultiSwarm
    G = len(gps)
    if len(cfgs) != G:
        raise ValueError("one PsoSettings per training set")
    traces = list(traces) if traces is not None else [None] * G
    target_costs = list(target_costs) if target_costs is not None else [None] * G
    records = list(records) if records is not None else [None] * G
    results: List[Optional[PsoResult]] = [None] * G
    runs: List[Optional[_SwarmRun]] = [None] * G
    for g in range(G):
        if cfgs[g].swarm_updater == "forced":
            results[g] = run_pso(gps[g], cfgs[g], traces[g], target_costs[g], records[g])
    multi = MultiSwarm(gps)
    try:
        for g in range(G):
            if results[g] is None:
                runs[g] = _SwarmRun(gps[g], cfgs[g], traces[g], target_costs[g], records[g])
                runs[g].start()
        while True:
            active = [g for g in range(G) if runs[g] is not None and not runs[g].done]
            if not active:
                break
            steps = {g: (runs[g].sw,) + runs[g].draw() for g in active}
            costs, infos = multi.pso_step(steps)
            for g in active:
                runs[g].after_step(costs[g], infos[g])
        for g in range(G):
            if runs[g] is not None:
                results[g] = runs[g].finish()
        return results
    finally:
        for r in runs:
            if r is not None:
                r.close()
        multi.close()


def _run_forced(gp, cfg: PsoSettings, trace, target_cost, record) -> PsoResult:
    """forced_swarm_update (pso.f90:485-558): the particles move one after
    another, each evaluated (get_cost with its re-randomisations,
    particle.f90:149-162) and selected before the next moves.  The state
    lives on the host and every particle is one evaluation of the library
    (a swarm of one), as the rule is sequential by definition."""
    S, P = cfg.swarm_size, gp.nparams
    rng = np.random.default_rng(cfg.seed)
    lo, hi = search_bounds(gp, cfg)
    res = PsoResult(best_position=np.empty(P), best_cost=math.inf)
    t0 = time.perf_counter()
    x = lo + rng.random((S, P)) * (hi - lo)
    v = (lo - hi) + rng.random((S, P)) * (2.0 * (hi - lo))
    ll, info = gp.swarm_loglik(x)
    cost = np.where(info == 0, -ll, math.inf)
    res.evaluations += S
    res.rejections += int(np.sum(info != 0))
    if trace:
        for k in range(S):
            trace(0, 0, k, int(info[k]), float(cost[k]), x[k])
    _record(record, it=0, kind="init", x=x, v=v, cost=cost, info=info)
    pbest, pcost = x.copy(), cost.copy()
    gcost, gbest = math.inf, x[0].copy()
    for k in range(S):                        # state%update in particle order
        if cost[k] <= gcost:
            gcost, gbest = float(cost[k]), x[k].copy()
    res.history.append((0, time.perf_counter() - t0, gcost))
    upd = ParameterUpdater(cfg.parameter_updater, cfg, rng, cost)
    _record(record, it=0, kind="coef", coef=list(upd.coef), draws=list(upd.draws))
    stall_cost, stall = math.inf, 0
    it = 1
    while True:
        for k in range(S):
            w, c1, c2 = upd.coef
            ra = rng.random(P)
            rb = rng.random(P)
            momentum = w * v[k]
            to_opt = c1 * ra * (pbest[k] - x[k])
            to_glob = c2 * rb * (gbest - x[k])
            v[k] = momentum + to_opt + to_glob
            partial = np.abs(v[k]) + np.abs(gbest - x[k])
            dv = None
            if np.any(partial < cfg.delta):
                dv = -cfg.delta + rng.random(P) * (cfg.delta - (-cfg.delta))
                v[k] = np.where(partial < cfg.delta, dv, v[k])
            x[k] = x[k] + v[k]
            rnd = 0
            while True:
                ll, inf = gp.swarm_loglik(x[k][None, :])
                res.evaluations += 1
                c = float(-ll[0]) if inf[0] == 0 else math.inf
                if trace:
                    trace(it, rnd, k, int(inf[0]), c, x[k])
                _record(record, it=it, kind="particle", k=k, round=rnd, coef=[w, c1, c2], r1=ra, r2=rb, dv=dv,
                        x=x[k], v=v[k], cost=c, info=int(inf[0]))
                if inf[0] == 0:
                    break
                res.rejections += 1
                rnd += 1
                if rnd > cfg.max_retries:
                    raise RuntimeError("PSO: particle stayed non-positive-definite after max_retries "
                                       "re-randomisations")
                x[k] = lo + rng.random(P) * (hi - lo)          # particle.f90:117-128
                v[k] = (lo - hi) + rng.random(P) * (2.0 * (hi - lo))
                dv = None
            cost[k] = c
            if c <= pcost[k]:
                pbest[k], pcost[k] = x[k].copy(), c
                if c <= gcost:
                    gcost, gbest = c, x[k].copy()
                    res.best_iter = it
            if upd.per_particle:
                upd.draws = []
                upd.update([c], first=k)
                _record(record, it=it, kind="coef", k=k, coef=list(upd.coef), draws=list(upd.draws))
        if upd.per_swarm:
            upd.draws = []
            upd.update(cost)
            _record(record, it=it, kind="coef", coef=list(upd.coef), draws=list(upd.draws))
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
    res.seconds = time.perf_counter() - t0
    res.best_cost, res.best_position = gcost, gbest
    return res
