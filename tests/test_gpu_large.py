"""Largest single-GPU workload: ibuprofen-like N=6000, d=93 (32 distances,
31 angles, 30 dihedrals), linear mean, concentrated likelihood, swarm 64
(64 x 6016^2 x 8 B = 18.5 GB of K resident).

* every particle: INFO consistent between the batched and plain paths, and
  their log-likelihoods bit-identical on a subset;
* two conditioned particles against LAPACK dpotrf / dpotrs on the oracle's
  K (the oracle's unblocked dpotf2 restatement takes minutes at this size;
  its kernel builder does not): relative error <= 1e-9, or 10 x the gap
  between two LAPACK orderings of the same K (K and its reversal) when
  that is larger.
"""
import numpy as np
import pytest
from scipy.linalg import lapack

import oracle
from ferebus_amd import SwarmLikelihood, compile_kernel, fit_mean, synthetic

pytestmark = pytest.mark.gpu


def _lapack_conc(k, y, n):
    c, info = lapack.dpotrf(k, lower=1)
    if info:
        return np.nan, int(info)
    logdet = 2.0 * np.sum(np.log(np.diag(c)))
    ry, _ = lapack.dpotrs(c, y, lower=1)
    r1, _ = lapack.dpotrs(c, np.ones(n), lower=1)
    r = y - ry.sum() / r1.sum()
    a, _ = lapack.dpotrs(c, r, lower=1)
    return -0.5 * n * np.log(r @ a / n) - 0.5 * logdet, 0


def test_ibuprofen_full_size(gpu):
    w = synthetic.ibuprofen(n=6000, seed=4)
    d = w.x.shape[1]
    prog = compile_kernel(w.kernel, w.kernels, d)
    mfit = fit_mean(w.mean, w.x, w.y)
    gp = SwarmLikelihood(w.x, w.y, prog, mean=mfit, noise=w.noise, likelihood=w.likelihood, max_swarm=64)
    rng = np.random.default_rng(9)
    th = rng.uniform(0.0, 1.0, (64, gp.nparams))
    th[:2, :] += 2.0          # short lengthscales: conditioned matrices for the LAPACK comparison
    ll, info = gp.swarm_loglik(th)
    print("ibuprofen INFO", info.tolist(), "path", gp.factor_path(64))
    assert np.isfinite(ll[info == 0]).all()
    pl, pinfo = gp.particle_loglik(th[:4])
    assert np.array_equal(pinfo, info[:4])
    assert np.array_equal(pl[pinfo == 0], ll[:4][pinfo == 0])
    spec = oracle.Spec.from_program(prog)
    n = len(w.y)
    for s in range(2):
        assert info[s] == 0
        k = oracle.kernel_matrix(w.x, spec, th[s])
        k = k + np.tril(k, -1).T + w.noise * np.eye(n)
        ref, linfo = _lapack_conc(k, w.y, n)
        rev, _ = _lapack_conc(k[::-1, ::-1].copy(), w.y[::-1].copy(), n)
        gap = abs(rev - ref) / abs(ref)
        rel = abs(ll[s] - ref) / abs(ref)
        print(f"particle {s}: gpu {ll[s]:.15e} lapack {ref:.15e} rel {rel:.2e} (orderings gap {gap:.2e})")
        assert linfo == 0 and rel <= max(1e-9, 10.0 * gap)


def test_ibuprofen_gradient_and_polish(gpu):
    """The same workload's adjoint gradient (concentrated likelihood, linear
    mean, 93 parameters) against central finite differences on a subset of
    the parameters, and one L-BFGS polish of a particle lowering its cost."""
    from test_gpu_gradient import FD_RTOL
    from ferebus_amd import lbfgs
    w = synthetic.ibuprofen(n=6000, seed=4)
    d = w.x.shape[1]
    prog = compile_kernel(w.kernel, w.kernels, d)
    gp = SwarmLikelihood(w.x, w.y, prog, mean=fit_mean(w.mean, w.x, w.y), noise=w.noise,
                         likelihood=w.likelihood, max_swarm=4)
    rng = np.random.default_rng(11)
    th = rng.uniform(0.0, 1.0, (1, gp.nparams)) + 2.0      # short lengthscales: conditioned
    ll, g, info = gp.swarm_grad(th)
    print("ibuprofen gradient info", info.tolist(), "ll", ll.tolist())
    assert info[0] == 0 and np.isfinite(g).all()
    t = th[0]
    for q in (0, 31, 62, 63, 92):
        h = max(1e-6 * abs(t[q]), 1e-7)
        tp, tm = t.copy(), t.copy()
        tp[q] += h
        tm[q] -= h
        fd = (gp.swarm_loglik(tp)[0][0] - gp.swarm_loglik(tm)[0][0]) / (2 * h)
        err = abs(fd - g[0, q]) / max(np.abs(g[0]).max(), 1e-300)
        print(f"ibuprofen param {q}: adjoint {g[0, q]:.6e} fd {fd:.6e} rel {err:.2e}")
        assert err <= FD_RTOL
    xp, fp, evals, accepted = lbfgs.polish(gp, t, lbfgs.LbfgsSettings(steps=2))
    print(f"polish: cost {-ll[0]:.10e} -> {fp:.10e} in {accepted} steps, {evals} evaluations")
    assert accepted >= 1 and fp < -ll[0]
    gp.close()
