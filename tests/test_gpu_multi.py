"""Several training sets of different N evaluated by one call
(fb_multi_swarm_loglik, the batched uneven case), all factorised in one
persistent launch (FB_MULTI_PERSISTENT=1), by each handle's own batched path
one after another (=0), or by whichever of the two the call's first
evaluation timed faster (unset): every particle's ll and INFO must be
bit-identical to its own handle's fb_swarm_loglik (same tile arithmetic per
matrix), with handles of different kernels, likelihoods and swarm sizes
mixed, and a skipped (s = 0) handle left untouched."""
import time

import numpy as np
import pytest

import oracle
from ferebus_amd import KernelConfig, MultiSwarm, SwarmLikelihood, compile_kernel, fit_mean, synthetic

pytestmark = pytest.mark.gpu


def _gp(n, d, seed, kernel="k1+k2", kernels=None, likelihood="marginal", mean="constant", S=8):
    rng = np.random.default_rng(seed)
    x = synthetic.alf_features(rng, n, d)
    y = synthetic.bump_target(rng, x, nbumps=5)
    prog = compile_kernel(kernel, kernels or synthetic.rbf_periodic_kernels(d), d)
    gp = SwarmLikelihood(x, y, prog, mean=fit_mean(mean, x, y), noise=1e-6, likelihood=likelihood, max_swarm=S)
    return gp, x, y, prog, rng


@pytest.mark.parametrize("schedule", ["1", "0", None])
def test_multi_matches_each_handle(gpu, monkeypatch, schedule):
    if schedule is None:
        monkeypatch.delenv("FB_MULTI_PERSISTENT", raising=False)
    else:
        monkeypatch.setenv("FB_MULTI_PERSISTENT", schedule)
    specs = [dict(n=150, d=6, seed=81), dict(n=333, d=6, seed=82, likelihood="concentrated"),
             dict(n=520, d=9, seed=83, kernel="k1*k2"), dict(n=64, d=6, seed=84, mean="linear")]
    gps, thetas = [], []
    for sp, S in zip(specs, (5, 3, 7, 2)):
        gp, x, y, prog, rng = _gp(**sp)
        gps.append(gp)
        thetas.append(rng.uniform(0.0, 1.0, (S, gp.nparams)))
    multi = MultiSwarm(gps)
    lls, infos = multi.swarm_loglik(thetas)
    for g, (gp, th) in enumerate(zip(gps, thetas)):
        ll, info = gp.swarm_loglik(th)
        print(f"gp {g}: n={gp.n} info {infos[g].tolist()}")
        assert np.array_equal(info, infos[g]) and np.array_equal(ll, lls[g])
    # one training set skipped, the others re-evaluated with new particles
    thetas2 = [t[::-1].copy() for t in thetas]
    thetas2[1] = None
    lls2, infos2 = multi.swarm_loglik(thetas2)
    assert len(lls2[1]) == 0
    for g in (0, 2, 3):
        ll, info = gps[g].swarm_loglik(thetas2[g])
        assert np.array_equal(info, infos2[g]) and np.array_equal(ll, lls2[g])
    # and against the oracle for the largest one
    gp, x, y, prog, _ = _gp(**specs[2])
    spec = oracle.Spec.from_program(prog)
    oll, oinfo = oracle.swarm_loglik(x, y, fit_mean("constant", x, y).value, spec, thetas[2][:2], noise=1e-6)
    np.testing.assert_allclose(lls[2][:2], oll, rtol=1e-9)
    multi.close()


def test_multi_glycine_uneven_full_size(gpu, monkeypatch):
    """Glycine-like: 10 GPs, d=24, N ~ U[1000, 4000], 40 particles each --
    400 Cholesky factorisations of mixed size per PSO iteration in one launch."""
    monkeypatch.setenv("FB_MULTI_PERSISTENT", "1")
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
    print(f"400 factorisations: one launch {t_multi * 1e3:.1f} ms, one launch per GP {t_seq * 1e3:.1f} ms, "
          f"failed {sum(int((i != 0).sum()) for i in infos)}")
    multi.close()
