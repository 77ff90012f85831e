"""The device-resident particle swarm (fb_pso_*): update, retries and
selection on the GPU, one graph replay per iteration.

* Replay: every position the device evaluated is recomputed on the host
  from the same RNG stream with the reference's rules (pso.f90:405-459,
  particle.f90:117-162), using the device's costs for the selections; the
  positions must agree to rounding (the device may contract the update into
  FMAs), and the gbest must be the one the rules pick.
* The costs at the traced positions are held to the oracle as in
  test_gpu_parity.
* Ties go to the later particle, failed particles cost +Inf and are
  retried, and the graph path equals the eager path bit for bit.
"""
import math

import numpy as np
import pytest

import oracle
from ferebus_amd import SwarmLikelihood, compile_kernel, fit_mean, synthetic
from ferebus_amd.pso import DeviceSwarm, PsoSettings, run_pso, search_bounds
from test_gpu_parity import RTOL, _check, _lapack, _oracle

pytestmark = pytest.mark.gpu


def _gp(n=80, d=6, seed=5, S=8, optimise_noise=False, noise=1e-6):
    rng = np.random.default_rng(seed)
    x = synthetic.alf_features(rng, n, d)
    y = synthetic.bump_target(rng, x, nbumps=4)
    prog = compile_kernel("k1+k2", synthetic.rbf_periodic_kernels(d), d)
    mfit = fit_mean("constant", x, y)
    gp = SwarmLikelihood(x, y, prog, mean=mfit, noise=noise, optimise_noise=optimise_noise, max_swarm=S)
    return x, y, prog, mfit.value, gp


def _host_replay(cfg, P, lo, hi, evals):
    """Positions and gbest of the reference's rules, from the same RNG draws
    as run_pso, with the device's costs."""
    S = cfg.swarm_size
    rng = np.random.default_rng(cfg.seed)
    by_key = {}
    for e in evals:
        by_key.setdefault((e[0], e[1]), {})[e[2]] = e
    x = lo + rng.random((S, P)) * (hi - lo)
    v = (lo - hi) + rng.random((S, P)) * (2.0 * (hi - lo))
    worst = 0.0
    got = by_key[(0, 0)]
    worst = max(worst, max(np.abs(got[k][5] - x[k]).max() for k in range(S)))
    cost = np.array([got[k][4] for k in range(S)])
    pbest, pcost = x.copy(), cost.copy()
    gcost, gbest = math.inf, None
    for k in range(S):
        if cost[k] <= gcost:
            gcost, gbest = cost[k], x[k].copy()
    it = 1
    while (it, 0) in by_key:
        r1, r2 = rng.random((S, P)), rng.random((S, P))
        v = cfg.inertia_weight * v + cfg.cognitive_learning_rate * r1 * (pbest - x) \
            + cfg.social_learning_rate * r2 * (gbest[None, :] - x)
        x = x + v
        got = by_key[(it, 0)]
        scale = np.maximum(np.abs(x), 1.0)
        worst = max(worst, max((np.abs(got[k][5] - x[k]) / scale[k]).max() for k in range(S)))
        x = np.array([got[k][5] for k in range(S)])       # continue from the device's bits
        cost = np.array([got[k][4] for k in range(S)])
        info = np.array([got[k][3] for k in range(S)])
        rnd = 0
        todo = np.flatnonzero(info != 0)
        while len(todo):
            rnd += 1
            xr = lo + rng.random((len(todo), P)) * (hi - lo)
            vr = (lo - hi) + rng.random((len(todo), P)) * (2.0 * (hi - lo))
            got = by_key[(it, rnd)]
            for r, k in enumerate(todo):
                assert np.array_equal(got[k][5], xr[r])
                x[k], v[k] = xr[r], vr[r]
                cost[k] = got[k][4]
                info[k] = got[k][3]
            todo = todo[info[todo] != 0]
        for k in range(S):
            if cost[k] <= pcost[k]:
                pbest[k], pcost[k] = x[k].copy(), cost[k]
        for k in range(S):
            if cost[k] <= gcost:
                gcost, gbest = cost[k], x[k].copy()
        it += 1
    return gcost, gbest, worst


@pytest.mark.parametrize("optimise_noise", [False, True])
def test_device_pso_replay(gpu, optimise_noise):
    x, y, prog, m, gp = _gp(S=8, optimise_noise=optimise_noise)
    cfg = PsoSettings(swarm_size=8, iterations=12, seed=3)
    evals = []
    res = run_pso(gp, cfg, trace=lambda *e: evals.append((e[0], e[1], e[2], e[3], e[4], np.array(e[5]))))
    lo, hi = search_bounds(gp, cfg)
    gcost, gbest, worst = _host_replay(cfg, gp.nparams, lo, hi, evals)
    print(f"{len(evals)} evaluations, {res.rejections} rejections, position agreement {worst:.2e}, "
          f"gbest cost {res.best_cost:.12e}")
    assert worst <= 1e-12
    assert res.best_cost == gcost
    np.testing.assert_array_equal(res.best_position, gbest)
    # the costs at every traced position against the oracle
    th = np.array([e[5] for e in evals])
    ll = -np.array([e[4] for e in evals])
    info = np.array([e[3] for e in evals])
    oll, oinfo = _oracle(x, y, m, prog, th, 1e-6, optimise_noise, "marginal")
    ok = info == 0
    ll[~ok] = 0.0
    _check(ll, info, oll, oinfo, RTOL, "device PSO trace",
           _lapack(x, y, m, prog, th, 1e-6, optimise_noise, "marginal"))


def test_ties_and_failures(gpu):
    x, y, prog, m, gp = _gp(S=6, optimise_noise=True)
    P = gp.nparams
    sw = DeviceSwarm(gp, 6, 0.789, 1.494, 1.494)
    rng = np.random.default_rng(0)
    x0 = rng.uniform(0.0, 1.0, (6, P))
    x0[:, -1] = 1e-6
    ll, info = gp.swarm_loglik(x0)
    best = int(np.argmax(np.where(info == 0, ll, -np.inf)))
    x0[4] = x0[best]                   # a tie with a later particle
    x0[1, -1] = -1.0                   # negative nugget: not positive definite
    cost, info = sw.init(x0, np.zeros((6, P)))
    assert info[1] != 0 and cost[1] == math.inf
    gcost, gbest, gidx = sw.select()
    want = max(k for k in range(6) if cost[k] == cost.min())
    assert gidx == want and want >= 4 and gcost == cost.min()
    np.testing.assert_array_equal(gbest, x0[want])
    # retry particle 1 at a valid position: a lower cost takes gbest
    cr, ir = sw.retry([1], x0[best][None, :], np.zeros((1, P)))
    assert ir[0] == 0 and cr[0] == cost[best]
    gcost2, _, gidx2 = sw.select()
    assert gcost2 == gcost and gidx2 == want      # particle 1 comes before the tie holder
    xs, vs = sw.positions()
    np.testing.assert_array_equal(xs[1], x0[best])
    # a step with r1 = r2 = 0 and v = 0 keeps every position
    c3, i3 = sw.step(np.zeros((6, P)), np.zeros((6, P)))
    xs3, _ = sw.positions()
    np.testing.assert_array_equal(xs3, xs)
    ok = (i3 == 0) & np.isfinite(c3)
    np.testing.assert_array_equal(c3[ok], np.where(np.arange(6) == 1, cr[0], cost)[ok])
    sw.close()


def test_graph_equals_eager(gpu, monkeypatch):
    cfg = PsoSettings(swarm_size=8, iterations=6, seed=11)
    *_, gp = _gp(S=8)
    a = run_pso(gp, cfg)
    monkeypatch.setenv("FB_GRAPH", "0")
    *_, gp2 = _gp(S=8)
    b = run_pso(gp2, cfg)
    assert a.best_cost == b.best_cost
    np.testing.assert_array_equal(a.best_position, b.best_position)
    assert [h[2] for h in a.history] == [h[2] for h in b.history]


def test_methanol_pso_iterations(gpu):
    """Full-size swarm (N=2000, d=12, S=50) for a few iterations: finite
    costs, gbest never increases, and its cost matches a fresh evaluation."""
    w = synthetic.methanol(n=2000, seed=0)
    prog = compile_kernel(w.kernel, w.kernels, 12)
    gp = SwarmLikelihood(w.x, w.y, prog, mean=fit_mean("constant", w.x, w.y), noise=1e-10, max_swarm=50)
    res = run_pso(gp, PsoSettings(swarm_size=50, iterations=4, seed=1))
    hist = [h[2] for h in res.history]
    print("gbest history", hist, "rejections", res.rejections, "path", gp.factor_path(50))
    assert all(b <= a for a, b in zip(hist, hist[1:]))
    ll, info = gp.swarm_loglik(res.best_position[None, :])
    assert info[0] == 0 and -ll[0] == res.best_cost
