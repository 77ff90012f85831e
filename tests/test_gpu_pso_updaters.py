"""The swarm and parameter updaters on the GPU, replayed against the
reference's rules with the oracle's costs.

For each updater the PSO runs with a record of everything the host drew
and read back.  The replay then checks, step by step from the recorded
state:
* every evaluated cost against the FP64 oracle (relative 1e-9; the
  nugget is 1e-6, a conditioned case);
* the positions the device produced: v = w v + c1 r1 (pbest - x) +
  c2 r2 (gbest - x), x += v with the coefficients in force
  (pso.f90:422-434; forced: one particle after another with the delta rule,
  pso.f90:513-530), pbest / gbest from the oracle's costs by the reference's
  selection rules (pso.f90:440-459; forced: gbest only when pbest improved,
  pso.f90:536-544);
* the coefficient sequence: changed only when the updater runs -- after
  every swarm update for the unified variants (pso.f90:563-565), after every
  particle of the forced updater for the individual ones (pso.f90:550-552),
  never for those under the global updater;
* the final gbest.
"""
import math

import numpy as np
import pytest

import oracle
from ferebus_amd import SwarmLikelihood, compile_kernel, fit_mean, synthetic
from ferebus_amd.pso import PsoSettings, run_pso

pytestmark = pytest.mark.gpu


def _setup():
    w = synthetic.water(n=150, seed=3)
    prog = compile_kernel(w.kernel, w.kernels, 3)
    m = fit_mean("constant", w.x, w.y)
    gp = SwarmLikelihood(w.x, w.y, prog, mean=m, noise=1e-6, max_swarm=8)
    spec = oracle.Spec.from_program(prog)
    ocost = lambda th: -oracle.swarm_loglik(w.x, w.y, m.value, spec, np.atleast_2d(th), noise=1e-6)[0]
    return gp, ocost


def _close(a, b, rtol=1e-13):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b)))


@pytest.mark.parametrize("updater", ["constant", "unified_adaptive", "individual_adaptive", "random_unified",
                                     "random_individual"])
def test_global_updater_replay(gpu, updater):
    gp, ocost = _setup()
    cfg = PsoSettings(swarm_size=8, iterations=6, seed=5, parameter_updater=updater)
    rec = []
    res = run_pso(gp, cfg, record=rec)
    init = rec[0]
    x, v = init["x"], init["v"]
    cost = np.where(init["info"] == 0, ocost(x), math.inf)
    np.testing.assert_allclose(init["cost"][init["info"] == 0], cost[init["info"] == 0], rtol=1e-9)
    pbest, pcost = x.copy(), cost.copy()
    gcost, gbest = math.inf, None
    for k in range(len(cost)):
        if cost[k] <= gcost:
            gcost, gbest = cost[k], x[k].copy()
    coef = rec[1]["coef"]
    if updater == "constant":
        assert coef == [cfg.inertia_weight, cfg.cognitive_learning_rate, cfg.social_learning_rate]
    n_coef_changes = 0
    i = 2
    while i < len(rec):
        st = rec[i]
        assert st["kind"] == "step" and st["coef"] == coef
        w_, c1, c2 = coef
        vx = w_ * v + c1 * st["r1"] * (pbest - x) + c2 * st["r2"] * (gbest[None, :] - x)
        assert _close(st["v"], vx) and _close(st["x"], x + vx), f"iteration {st['it']}: update"
        x, v = st["x"].copy(), st["v"].copy()
        cost = np.where(st["info"] == 0, ocost(x), math.inf)
        ok = st["info"] == 0
        np.testing.assert_allclose(st["cost"][ok], cost[ok], rtol=1e-9)
        i += 1
        while i < len(rec) and rec[i]["kind"] == "retry":
            r = rec[i]
            x[r["idx"]], v[r["idx"]] = r["x"], r["v"]
            c = np.where(r["info"] == 0, ocost(r["x"]), math.inf)
            np.testing.assert_allclose(r["cost"][r["info"] == 0], c[r["info"] == 0], rtol=1e-9)
            cost[r["idx"]] = c
            i += 1
        for k in range(len(cost)):            # pbest (pso.f90:440-443), gbest (:454-459)
            if cost[k] <= pcost[k]:
                pcost[k], pbest[k] = cost[k], x[k].copy()
        for k in range(len(cost)):
            if cost[k] <= gcost:
                gcost, gbest = cost[k], x[k].copy()
        if i < len(rec) and rec[i]["kind"] == "coef":
            assert updater in ("unified_adaptive", "random_unified")
            coef = rec[i]["coef"]
            n_coef_changes += 1
            i += 1
    if updater in ("unified_adaptive", "random_unified"):
        assert n_coef_changes == cfg.iterations
    else:
        assert n_coef_changes == 0
    print(f"{updater}: replayed gbest {gcost:.10f}, device {res.best_cost:.10f}")
    assert abs(res.best_cost - gcost) <= 1e-9 * abs(gcost)
    assert _close(res.best_position, gbest, 1e-12)


@pytest.mark.parametrize("updater,delta", [("constant", 0.3), ("unified_adaptive", 1e-3),
                                           ("individual_adaptive", 1e-3), ("random_individual", 0.05)])
def test_forced_updater_replay(gpu, updater, delta):
    gp, ocost = _setup()
    cfg = PsoSettings(swarm_size=6, iterations=4, seed=9, swarm_updater="forced", parameter_updater=updater,
                      delta=delta)
    rec = []
    res = run_pso(gp, cfg, record=rec)
    init = rec[0]
    x, v = init["x"].copy(), init["v"].copy()
    cost = np.where(init["info"] == 0, ocost(x), math.inf)
    pbest, pcost = x.copy(), cost.copy()
    gcost, gbest = math.inf, None
    for k in range(len(cost)):
        if cost[k] <= gcost:
            gcost, gbest = cost[k], x[k].copy()
    coef = rec[1]["coef"]
    n_redraw = n_particle_coef = n_swarm_coef = 0
    for e in rec[2:]:
        if e["kind"] == "coef":
            if "k" in e:
                assert updater in ("individual_adaptive", "random_individual")
                n_particle_coef += 1
            else:
                assert updater == "unified_adaptive"
                n_swarm_coef += 1
            coef = e["coef"]
            continue
        assert e["kind"] == "particle"
        k = e["k"]
        if e["round"] == 0:
            assert e["coef"] == coef
            w_, c1, c2 = coef
            vk = w_ * v[k] + c1 * e["r1"] * (pbest[k] - x[k]) + c2 * e["r2"] * (gbest - x[k])
            pp = np.abs(vk) + np.abs(gbest - x[k])
            if np.any(pp < cfg.delta):
                assert e["dv"] is not None and np.all(np.abs(e["dv"]) <= cfg.delta)
                vk = np.where(pp < cfg.delta, e["dv"], vk)
                n_redraw += 1
            else:
                assert e["dv"] is None
            assert _close(e["v"], vk) and _close(e["x"], x[k] + vk), f"iteration {e['it']} particle {k}"
        x[k], v[k] = e["x"], e["v"]
        c = ocost(x[k])[0] if e["info"] == 0 else math.inf
        if e["info"] == 0:
            np.testing.assert_allclose(e["cost"], c, rtol=1e-9)
            if c <= pcost[k]:
                pbest[k], pcost[k] = x[k].copy(), c
                if c <= gcost:
                    gcost, gbest = c, x[k].copy()
    if delta >= 0.3:
        assert n_redraw > 0               # the delta rule was exercised
    if updater in ("individual_adaptive", "random_individual"):
        assert n_particle_coef == cfg.iterations * cfg.swarm_size
    if updater == "unified_adaptive":
        assert n_swarm_coef == cfg.iterations
    print(f"forced {updater}: {n_redraw} delta re-draws, replayed gbest {gcost:.10f}, device {res.best_cost:.10f}")
    assert abs(res.best_cost - gcost) <= 1e-9 * abs(gcost)
    assert _close(res.best_position, gbest, 1e-12)
