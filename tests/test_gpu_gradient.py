"""Adjoint gradient d logL / d theta (fb_swarm_loglik_grad) through the C ABI.

Checked two independent ways:
* against the oracle's FP64 CPU gradient, which forms W = alpha alpha^T / s2 -
  K^{-1} densely (K^{-1} column by column with dpotrs) and differentiates
  every leaf of the kernel tree; relative error, per particle, scaled by the
  largest gradient entry, <= 1e-8.  One exception: a linear leaf adds
  c 11^T to K, and its derivative needs 1^T K^{-1} 1, which any method that
  forms K^{-1} (the oracle's included) computes by cancellation of N^2 large
  entries; at the 1e-6 nugget that case is held to 1e-6 against the oracle
  (measured 1.5e-8), and a well-conditioned linear case (nugget 1e-2) is held
  to the full 1e-8 and to finite differences;
* against central finite differences of the GPU's own log-likelihood,
  h = 1e-6 * |theta| (floor 1e-7), relative error <= 1e-5 (the FD truncation
  and rounding error at these step sizes is ~1e-7 of the gradient scale).
The gradient also returns the log-likelihood itself, which must equal the
plain swarm evaluation bit for bit.
"""
import numpy as np
import pytest

import oracle
from ferebus_amd import KernelConfig, SwarmLikelihood, compile_kernel, fit_mean, synthetic

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-8
FD_RTOL = 1e-5

CASES = {
    "rbf+periodic": ("k1+k2", None),
    "product": ("k1*k2", None),
    "scaled": ("2*k1", {"k1": KernelConfig("k1", "rbf", [1, 2, 3])}),
    "nested": ("(k1+k2)*k3", {"k1": KernelConfig("k1", "rbf", [1, 2]), "k2": KernelConfig("k2", "rbf", [3, 4]),
                              "k3": KernelConfig("k3", "periodic", [6])}),
    "linear": ("k1+k2", {"k1": KernelConfig("k1", "rbf", [1, 2, 4]), "k2": KernelConfig("k2", "linear", [3, 5])}),
    "cyclic": ("k1", {"k1": KernelConfig("k1", "rbf-cyclic")}),
    "cyclic_sum": ("k1+k2", {"k1": KernelConfig("k1", "rbf-cyclic"), "k2": KernelConfig("k2", "periodic", [3, 6])}),
}


def _setup(n, d, seed, kernel="k1+k2", kernels=None, likelihood="marginal", mean="constant", noise=1e-6,
           optimise_noise=False, S=3):
    rng = np.random.default_rng(seed)
    x = synthetic.alf_features(rng, n, d)
    y = synthetic.bump_target(rng, x, nbumps=5)
    prog = compile_kernel(kernel, kernels or synthetic.rbf_periodic_kernels(d), d)
    mfit = fit_mean(mean, x, y)
    gp = SwarmLikelihood(x, y, prog, mean=mfit, noise=noise, optimise_noise=optimise_noise,
                         likelihood=likelihood, max_swarm=max(S, 8))
    th = rng.uniform(0.0, 1.0, (S, gp.nparams))
    return x, y, prog, mfit.value, gp, th


def _oracle_grad(x, y, m, prog, th, noise, optimise_noise, likelihood):
    spec = oracle.Spec.from_program(prog)
    out = [oracle.loglik_grad(x, y, m, spec, t, noise=noise, optimise_noise=optimise_noise,
                              likelihood=likelihood) for t in th]
    return (np.array([o[0] for o in out]), np.array([o[1] for o in out]),
            np.array([o[2] for o in out]))


def _rel(g, ref):
    scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
    return np.max(np.abs(g - ref) / scale, axis=1)


def _check_against_oracle(x, y, m, prog, gp, th, noise, optimise_noise, likelihood, label, rtol=GRAD_RTOL):
    ll, g, info = gp.swarm_grad(th)
    oll, og, oinfo = _oracle_grad(x, y, m, prog, th, noise, optimise_noise, likelihood)
    print(f"{label}: info {info.tolist()} oracle {oinfo.tolist()}")
    assert np.array_equal(info, oinfo)
    ok = info == 0
    rel = _rel(g[ok], og[ok])
    print(f"{label}: max relative gradient error {rel.max():.3e}")
    assert (rel <= rtol).all(), f"{label}: gradient error {rel.max():.3e}"
    # the returned ll is the swarm evaluation's
    sl, sinfo = gp.swarm_loglik(th)
    assert np.array_equal(sinfo, info) and np.array_equal(sl[ok], ll[ok])
    np.testing.assert_allclose(ll[ok], oll[ok], rtol=1e-9)
    return g


@pytest.mark.parametrize("n", [64, 150, 200])
def test_gradient_sizes(gpu, n):
    x, y, prog, m, gp, th = _setup(n, 6, seed=21 + n)
    _check_against_oracle(x, y, m, prog, gp, th, 1e-6, False, "marginal", f"grad n={n}")


@pytest.mark.parametrize("name", sorted(CASES))
def test_gradient_kernel_compositions(gpu, name):
    expr, ks = CASES[name]
    x, y, prog, m, gp, th = _setup(140, 6, seed=31, kernel=expr, kernels=ks)
    if name == "scaled":
        th[:, 0] = np.abs(th[:, 0]) + 0.5
    _check_against_oracle(x, y, m, prog, gp, th, 1e-6, False, "marginal", f"grad {name}",
                          rtol=1e-6 if name == "linear" else GRAD_RTOL)


def test_gradient_linear_well_conditioned(gpu):
    expr, ks = CASES["linear"]
    x, y, prog, m, gp, th = _setup(140, 6, seed=33, kernel=expr, kernels=ks, noise=1e-2)
    g = _check_against_oracle(x, y, m, prog, gp, th, 1e-2, False, "marginal", "grad linear nugget 1e-2")
    fd = _fd(gp, th[0])
    rel = _rel(g[:1], fd[None, :])
    print(f"grad linear (nugget 1e-2) vs central FD: {rel.max():.3e}")
    assert rel.max() <= FD_RTOL


def test_gradient_optimised_noise(gpu):
    x, y, prog, m, gp, th = _setup(130, 6, seed=41, optimise_noise=True)
    th[:, -1] = [1e-5, 1e-4, 1e-3]
    g = _check_against_oracle(x, y, m, prog, gp, th, 1e-6, True, "marginal", "grad optimise_noise")
    assert np.all(np.isfinite(g[:, -1]))


@pytest.mark.parametrize("mean", ["zero", "linear"])
def test_gradient_concentrated(gpu, mean):
    x, y, prog, m, gp, th = _setup(160, 6, seed=51, likelihood="concentrated", mean=mean)
    _check_against_oracle(x, y, m, prog, gp, th, 1e-6, False, "concentrated", f"grad concentrated {mean}")


def _fd(gp, t):
    g = np.zeros_like(t)
    for q in range(len(t)):
        h = max(1e-6 * abs(t[q]), 1e-7)
        tp, tm = t.copy(), t.copy()
        tp[q] += h
        tm[q] -= h
        lp, ip = gp.swarm_loglik(tp)
        lm, im = gp.swarm_loglik(tm)
        assert ip[0] == 0 and im[0] == 0
        g[q] = (lp[0] - lm[0]) / (2 * h)
    return g


def test_gradient_finite_differences(gpu):
    x, y, prog, m, gp, th = _setup(200, 6, seed=61, optimise_noise=True, S=1)
    th[0, -1] = 1e-4
    ll, g, info = gp.swarm_grad(th)
    assert info[0] == 0
    fd = _fd(gp, th[0])
    rel = _rel(g, fd[None, :])
    print(f"grad vs central FD: max relative error {rel.max():.3e}")
    assert rel.max() <= FD_RTOL


def test_gradient_nma_full_size(gpu):
    """NMA-like N=3000, d=30: ARD rbf (30) + periodic (9 dihedrals) + optimised
    noise, marginal likelihood -- the gradient workload at full size, checked
    by central finite differences on a subset of the 40 parameters."""
    w = synthetic.nma(n=3000, seed=3)
    prog = compile_kernel(w.kernel, w.kernels, w.x.shape[1])
    gp = SwarmLikelihood(w.x, w.y, prog, mean=fit_mean(w.mean, w.x, w.y), noise=w.noise,
                         optimise_noise=True, likelihood=w.likelihood, max_swarm=2)
    rng = np.random.default_rng(5)
    th = rng.uniform(0.0, 1.0, (2, gp.nparams))
    th[:, -1] = [1e-5, 1e-4]
    ll, g, info = gp.swarm_grad(th)
    print("nma info", info.tolist(), "ll", ll.tolist())
    assert (info == 0).all() and np.isfinite(g).all()
    t = th[1]
    for q in (0, 7, 29, 30, 35, 39):
        h = max(1e-6 * abs(t[q]), 1e-7)
        tp, tm = t.copy(), t.copy()
        tp[q] += h
        tm[q] -= h
        fd = (gp.swarm_loglik(tp)[0][0] - gp.swarm_loglik(tm)[0][0]) / (2 * h)
        err = abs(fd - g[1, q]) / max(np.abs(g[1]).max(), 1e-300)
        print(f"nma param {q}: adjoint {g[1, q]:.6e} fd {fd:.6e} rel {err:.2e}")
        assert err <= FD_RTOL


def test_lbfgs_polish_improves(gpu):
    """ferebus_amd.lbfgs (mirror of src/optimiser/fb_lbfgs.f90) on the adjoint
    gradient: every accepted step lowers the cost, the result is
    reproducible, and the final cost is the swarm evaluation's at that theta."""
    from ferebus_amd import lbfgs
    x, y, prog, m, gp, th = _setup(180, 6, seed=71, S=1)
    ll0, _ = gp.swarm_loglik(th)
    xp, fp, evals, accepted = lbfgs.polish(gp, th[0], lbfgs.LbfgsSettings(steps=6))
    print(f"lbfgs: cost {-ll0[0]:.10e} -> {fp:.10e} in {accepted} steps, {evals} evaluations")
    assert accepted >= 1 and fp < -ll0[0]
    xp2, fp2, _, _ = lbfgs.polish(gp, th[0], lbfgs.LbfgsSettings(steps=6))
    assert np.array_equal(xp, xp2) and fp == fp2
    ll, info = gp.swarm_loglik(xp)
    assert info[0] == 0 and -ll[0] == fp


def test_lbfgs_batched_trials_match_sequential(gpu):
    """Trial steps evaluated four at a time accept the same step as the
    one-at-a-time backtracking search."""
    from ferebus_amd import lbfgs
    x, y, prog, m, gp, th = _setup(180, 6, seed=71, S=4)
    a = lbfgs.polish(gp, th[0], lbfgs.LbfgsSettings(steps=6, batch=1))
    b = lbfgs.polish(gp, th[0], lbfgs.LbfgsSettings(steps=6, batch=4))
    print(f"sequential: cost {a[1]:.12e}, {a[3]} steps, {a[2]} evaluations; "
          f"batched: cost {b[1]:.12e}, {b[3]} steps, {b[2]} evaluations")
    assert a[3] == b[3]
    np.testing.assert_allclose(b[0], a[0], rtol=1e-12, atol=0.0)
    assert abs(b[1] - a[1]) <= 1e-12 * abs(a[1])
