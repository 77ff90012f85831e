"""Several training sets of different N evaluated by one call
(fb_multi_swarm_loglik, the batched uneven case), by each handle's own
batched path one after another (FB_MULTI_SCHEDULE=0), all factorised in one
persistent launch (=1), by the chained lock-step schedule with every
handle's matrices in each launch, replayed as one hipGraph (=2), or by
whichever of them the call's first evaluation timed fastest (unset): every
particle's ll and INFO must be bit-identical to its own handle's
fb_swarm_loglik (same tile arithmetic per matrix), with handles of different
kernels, likelihoods and swarm sizes mixed, and a skipped (s = 0) handle
left untouched."""
import time

import numpy as np
import pytest

import oracle
from ferebus_amd import KernelConfig, MultiSwarm, SwarmLikelihood, compile_kernel, fit_mean, synthetic
from ferebus_amd.pso import PsoSettings, run_pso, run_pso_multi
from test_gpu_parity import RTOL, _check, _lapack, _oracle

pytestmark = pytest.mark.gpu


def _gp(n, d, seed, kernel="k1+k2", kernels=None, likelihood="marginal", mean="constant", S=8):
    rng = np.random.default_rng(seed)
    x = synthetic.alf_features(rng, n, d)
    y = synthetic.bump_target(rng, x, nbumps=5)
    prog = compile_kernel(kernel, kernels or synthetic.rbf_periodic_kernels(d), d)
    gp = SwarmLikelihood(x, y, prog, mean=fit_mean(mean, x, y), noise=1e-6, likelihood=likelihood, max_swarm=S)
    return gp, x, y, prog, rng


def _schedule(monkeypatch, schedule, tile=None, tail=None):
    monkeypatch.delenv("FB_MULTI_PERSISTENT", raising=False)
    if schedule is None:
        monkeypatch.delenv("FB_MULTI_SCHEDULE", raising=False)
    else:
        monkeypatch.setenv("FB_MULTI_SCHEDULE", schedule)
    for key, v in (("FB_MULTI_TILE", tile), ("FB_MULTI_TAIL", tail)):
        if v is None:
            monkeypatch.delenv(key, raising=False)
        else:
            monkeypatch.setenv(key, str(v))


SPECS = [dict(n=150, d=6, seed=81), dict(n=333, d=6, seed=82, likelihood="concentrated"),
         dict(n=520, d=9, seed=83, kernel="k1*k2"), dict(n=64, d=6, seed=84, mean="linear")]


def _mixed(sizes=(5, 3, 7, 2)):
    gps, thetas = [], []
    for sp, S in zip(SPECS, sizes):
        gp, x, y, prog, rng = _gp(**sp)
        gps.append(gp)
        thetas.append(rng.uniform(0.0, 1.0, (S, gp.nparams)))
    return gps, thetas


def _same_as_handles(gps, thetas, lls, infos):
    for g, (gp, th) in enumerate(zip(gps, thetas)):
        if th is None:
            assert len(lls[g]) == 0
            continue
        ll, info = gp.swarm_loglik(th)
        print(f"gp {g}: n={gp.n} info {infos[g].tolist()}")
        assert np.array_equal(info, infos[g]) and np.array_equal(ll, lls[g]), f"gp {g}"


@pytest.mark.parametrize("schedule", ["1", "0", "2", None])
def test_multi_matches_each_handle(gpu, monkeypatch, schedule):
    _schedule(monkeypatch, schedule)
    gps, thetas = _mixed()
    multi = MultiSwarm(gps)
    lls, infos = multi.swarm_loglik(thetas)
    chosen = multi.schedule()
    print("schedule:", chosen)
    assert chosen == (MultiSwarm.SCHEDULES[int(schedule)] if schedule is not None else chosen) and chosen
    _same_as_handles(gps, thetas, lls, infos)
    # the same counts with new particles (the merged schedule replays its graph)
    thetas2 = [t[::-1].copy() for t in thetas]
    lls2, infos2 = multi.swarm_loglik(thetas2)
    _same_as_handles(gps, thetas2, lls2, infos2)
    # one training set skipped, the others re-evaluated
    thetas3 = [t[1:].copy() for t in thetas]
    thetas3[1] = None
    lls3, infos3 = multi.swarm_loglik(thetas3)
    _same_as_handles(gps, thetas3, lls3, infos3)
    # and against the oracle for the largest one
    gp, x, y, prog, _ = _gp(**SPECS[2])
    spec = oracle.Spec.from_program(prog)
    oll, oinfo = oracle.swarm_loglik(x, y, fit_mean("constant", x, y).value, spec, thetas[2][:2], noise=1e-6)
    np.testing.assert_allclose(lls[2][:2], oll, rtol=1e-9)
    multi.close()


@pytest.mark.parametrize("tile,tail", [(0, 3), (1, 3), (2, 0), (4, 1), (8, 0), (8, 5), (12, 2), (16, 3), (20, 4)])
def test_merged_tile_codes(gpu, monkeypatch, tile, tail):
    """The merged schedule with each panel tile code (row-tile shapes, two
    role work-groups, fused row solves, 8-deep three-per-CU) and a number of
    last panels pre-updated in the launch before them: the bits of each
    handle's own evaluation."""
    _schedule(monkeypatch, "2", tile, tail)
    gps, thetas = _mixed()
    multi = MultiSwarm(gps)
    lls, infos = multi.swarm_loglik(thetas)
    _same_as_handles(gps, thetas, lls, infos)
    multi.close()


def test_merged_more_segments_than_one_launch(gpu, monkeypatch):
    """19 training sets: more than a launch's 16 segments, so each panel
    step takes two launches; the bits of each handle's own evaluation."""
    _schedule(monkeypatch, "2")
    gps, thetas = [], []
    for g in range(19):
        gp, x, y, prog, rng = _gp(n=60 + 23 * g, d=6, seed=200 + g, S=3)
        gps.append(gp)
        thetas.append(rng.uniform(0.0, 1.0, (1 + g % 3, gp.nparams)))
    multi = MultiSwarm(gps)
    lls, infos = multi.swarm_loglik(thetas)
    _same_as_handles(gps, thetas, lls, infos)
    multi.close()


def test_merged_wait_bound_reported(gpu, monkeypatch):
    """A hand-over wait of the merged launches that reaches its bound ends
    the kernel normally and the call returns an error (forced from the host
    with a bound of 0 ticks)."""
    _schedule(monkeypatch, "2")
    monkeypatch.setenv("FB_CHAIN_WAIT_TICKS", "0")
    gps, thetas = _mixed()
    multi = MultiSwarm(gps)
    with pytest.raises(RuntimeError, match="bound"):
        multi.swarm_loglik(thetas)
    multi.close()
    for gp in gps:
        gp.close()


def _same_record(a, b, label):
    assert len(a) == len(b), label
    for ea, eb in zip(a, b):
        assert ea.keys() == eb.keys(), label
        for k in ea:
            va, vb = ea[k], eb[k]
            if isinstance(va, np.ndarray):
                assert np.array_equal(va, vb), (label, ea["it"], ea["kind"], k)
            else:
                assert va == vb, (label, ea["it"], ea["kind"], k)


@pytest.mark.parametrize("schedule", ["2", "0", None])
def test_multi_pso_matches_each_run(gpu, monkeypatch, schedule):
    """run_pso_multi over four training sets (different N, kernels,
    likelihoods, swarm sizes, parameter updaters; one stopping early on the
    relative change, one with the forced updater run on its own) gives each
    the run of its own run_pso, bit for bit: every drawn coefficient and
    uniform, position, velocity, cost and INFO of every iteration, and the
    final gbest.  The schedules: merged (2), sequential (0), as timed."""
    _schedule(monkeypatch, schedule)
    gps = [_gp(**sp, S=12)[0] for sp in SPECS] + [_gp(n=90, d=6, seed=85, S=12)[0]]
    cfgs = [PsoSettings(swarm_size=10, iterations=7, seed=40),
            PsoSettings(swarm_size=6, iterations=7, seed=41, parameter_updater="unified_adaptive"),
            PsoSettings(swarm_size=12, iterations=40, seed=42, relative_change=True, stall_iterations=2,
                        tolerance=1e-2),
            PsoSettings(swarm_size=4, iterations=9, seed=43, parameter_updater="random_unified"),
            PsoSettings(swarm_size=5, iterations=3, seed=44, swarm_updater="forced")]
    rec_m = [[] for _ in gps]
    res_m = run_pso_multi(gps, cfgs, records=rec_m)
    for g, (gp, cfg) in enumerate(zip(gps, cfgs)):
        rec = []
        res = run_pso(gp, cfg, record=rec)
        print(f"gp {g}: n={gp.n} iterations {res.iterations} / {res_m[g].iterations} cost {res.best_cost}")
        assert res.iterations == res_m[g].iterations and res.evaluations == res_m[g].evaluations
        assert res.best_cost == res_m[g].best_cost
        assert np.array_equal(res.best_position, res_m[g].best_position)
        assert [h[2] for h in res.history] == [h[2] for h in res_m[g].history]
        _same_record(rec, rec_m[g], f"gp {g}")
    assert res_m[2].iterations < 40      # stopped on the relative change while others ran on


@pytest.mark.parametrize("schedule", ["2", None])
def test_multi_glycine_uneven_full_size(gpu, monkeypatch, schedule):
    """Glycine-like: 10 GPs, d=24, N ~ U[1000, 4000], 40 particles each --
    400 Cholesky factorisations of mixed size per PSO iteration in one call;
    every GP bit-identical to its own evaluation, the smallest checked
    against the oracle."""
    _schedule(monkeypatch, schedule)
    ws = synthetic.glycine_gps(seed=2)
    gps, thetas = [], []
    rng = np.random.default_rng(9)
    for w in ws:
        prog = compile_kernel(w.kernel, w.kernels, w.x.shape[1])
        gp = SwarmLikelihood(w.x, w.y, prog, mean=fit_mean(w.mean, w.x, w.y), noise=w.noise,
                             likelihood=w.likelihood, max_swarm=40)
        gps.append(gp)
        thetas.append(rng.uniform(0.0, 1.0, (40, gp.nparams)))
    print("N per GP:", [gp.n for gp in gps])
    multi = MultiSwarm(gps)
    multi.swarm_loglik(thetas)                      # first call builds the schedule
    t0 = time.perf_counter()
    lls, infos = multi.swarm_loglik(thetas)
    t_multi = time.perf_counter() - t0
    t_seq = 0.0
    for g, (gp, th) in enumerate(zip(gps, thetas)):
        gp.swarm_loglik(th)
        t0 = time.perf_counter()
        ll, info = gp.swarm_loglik(th)
        t_seq += time.perf_counter() - t0
        assert np.array_equal(info, infos[g]) and np.array_equal(ll, lls[g]), f"gp {g}"
    print(f"400 factorisations: one call {t_multi * 1e3:.1f} ms, one call per GP {t_seq * 1e3:.1f} ms, "
          f"failed {sum(int((i != 0).sum()) for i in infos)}")
    # the smallest GP against the oracle (1e-10 nugget: the tolerance takes
    # in the oracle's own gap to LAPACK on the same K)
    g = int(np.argmin([gp.n for gp in gps]))
    w = ws[g]
    prog = compile_kernel(w.kernel, w.kernels, w.x.shape[1])
    m = fit_mean(w.mean, w.x, w.y).value
    th = thetas[g][:2]
    oll, oinfo = _oracle(w.x, w.y, m, prog, th, w.noise, False, w.likelihood)
    _check(lls[g][:2], infos[g][:2], oll, oinfo, RTOL, f"glycine gp {g} n={w.x.shape[0]}",
           _lapack(w.x, w.y, m, prog, th, w.noise, False, w.likelihood))
    multi.close()
