"""Pins of the CPU oracle (oracle/oracle.c) -- runs without a GPU.

Each oracle function is checked against something that does not restate it:
* golden vectors written with scipy's LAPACK dpotrf/dpotrs and BLAS dgemv on
  the reference's tensor formulation (tests/golden/make_golden.py);
* closed forms: K = a 11' + b I (Sherman-Morrison determinant and solve,
  concentrated mu_hat = mean(y), sigma2_hat = var(y)/b) and its LAPACK INFO;
* special values of the kernels (a full period of the periodic kernel,
  Delta = pi, vanishing lengthscales), kernel symmetry;
* extended-precision brute force on tiny inputs;
* numpy least squares for the linear / quadratic means;
* central finite differences for the dense gradient on a conditioned case.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from ferebus_amd.kernel_expr import (LEAF_CONSTANT, LEAF_LINEAR, LEAF_PERIODIC, LEAF_RBF,
                                     LEAF_RBF_CYCLIC, OP_PROD, OP_SUM)

GOLDEN = Path(__file__).resolve().parent / "golden" / "golden_lapack.npz"
KIND = {"rbf": LEAF_RBF, "periodic": LEAF_PERIODIC, "rbf-cyclic": LEAF_RBF_CYCLIC,
        "linear": LEAF_LINEAR, "constant": LEAF_CONSTANT}


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.load()


def spec_from_golden(program):
    nt, nd, ad = [], [], []
    for node in program:
        if node[0] == "op":
            nt.append(OP_SUM if node[1] == "+" else OP_PROD)
            nd.append(0)
        else:
            nt.append(KIND[node[1]])
            nd.append(0 if node[1] == "constant" else len(node[2]))
            ad.extend(node[2])
    return oracle.Spec(nt, nd, ad)


def golden_cases():
    z = np.load(GOLDEN)
    meta = json.loads(bytes(z["meta"]).decode())
    for k, m in enumerate(meta):
        yield m, {key: z[f"{k}_{key}"] for key in ("x", "y", "theta", "mean", "alpha", "out4")}


def test_golden_vectors_lapack():
    n_checked = 0
    for m, a in golden_cases():
        spec = spec_from_golden(m["program"])
        ll, info, alpha, out4 = oracle.loglik(a["x"], a["y"], a["mean"], spec, a["theta"], noise=m["noise"],
                                              optimise_noise=m["optimise_noise"], likelihood=m["likelihood"])
        assert info == m["info"], m["name"]
        if info == 0:
            # two backward-stable Cholesky orderings differ by ~cond(K) eps
            tol = max(1e-10, 10.0 * m["cond"] * np.finfo(float).eps)
            assert abs(ll - m["ll"]) <= tol * abs(m["ll"]), (m["name"], ll, m["ll"], m["cond"])
            np.testing.assert_allclose(alpha, a["alpha"], rtol=0, atol=tol * 1e3 * np.abs(a["alpha"]).max())
            np.testing.assert_allclose(out4[1:], a["out4"][1:], rtol=tol, atol=1e-12)
        n_checked += 1
    assert n_checked >= 10


def test_golden_means():
    for m, a in golden_cases():
        mv, _, _, _, info = oracle.mean(m["mean"], a["x"], a["y"])
        assert info == 0
        np.testing.assert_allclose(mv, a["mean"], rtol=1e-12, atol=1e-13)


def test_means_vs_numpy_lstsq():
    rng = np.random.default_rng(3)
    x = rng.uniform(1.5, 3.0, (60, 4))
    y = 0.3 * x[:, 0] - 0.7 * x[:, 2] + 0.05 * rng.normal(size=60)
    for kind, order in (("linear", 1), ("quadratic", 2)):
        m, beta, xmin, ymin, info = oracle.mean(kind, x, y)
        imin = int(np.argmin(y))
        a = (x - x[imin]) ** order
        b_ref = np.linalg.lstsq(a, y - y[imin], rcond=None)[0]
        np.testing.assert_allclose(beta, b_ref, rtol=1e-9)
        np.testing.assert_allclose(xmin, x[imin])
        assert ymin == y[imin]
        np.testing.assert_allclose(m, a @ b_ref + y[imin], rtol=1e-12)
    m, *_ = oracle.mean("constant", x, y)
    assert np.allclose(m, y.mean())
    m, *_ = oracle.mean("zero", x, y)
    assert not m.any()


def _constant_spec():
    return oracle.Spec([LEAF_CONSTANT], [0], [])


@pytest.mark.parametrize("n,a,b", [(1, 0.7, 0.3), (17, 2.0, 0.5), (64, 0.25, 1e-3), (130, 1.0, 1.0)])
def test_closed_form_a11_plus_bI(n, a, b):
    """K = a 11' + b I built by the constant kernel (value theta = a) + nugget b."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 2))
    y = rng.normal(size=n) + 3.0
    spec = _constant_spec()
    ll, info, alpha, out4 = oracle.loglik(x, y, None, spec, [a], noise=b)
    assert info == 0
    logdet = (n - 1) * np.log(b) + np.log(b + n * a)
    quad = (y @ y - a * y.sum() ** 2 / (b + n * a)) / b
    np.testing.assert_allclose(out4[1], logdet, rtol=1e-12)
    np.testing.assert_allclose(ll, -0.5 * quad - 0.5 * logdet - 0.5 * n * np.log(2 * np.pi), rtol=1e-11)
    alpha_ref = (y - a * y.sum() / (b + n * a)) / b
    np.testing.assert_allclose(alpha, alpha_ref, rtol=1e-10, atol=1e-12)
    if n > 1:
        llc, infoc, _, o4c = oracle.loglik(x, y, None, spec, [a], noise=b, likelihood="concentrated")
        np.testing.assert_allclose(o4c[2], y.mean(), rtol=1e-11)
        np.testing.assert_allclose(o4c[3], np.var(y) / b, rtol=1e-10)
        np.testing.assert_allclose(llc, -0.5 * n * np.log(np.var(y) / b) - 0.5 * logdet, rtol=1e-11)


def test_closed_form_info():
    """Pivots of a 11' + b I are (a+b), b(2a+b)/(a+b), ... : b < 0 < a+b fails
    at column 2; a + b <= 0 fails at column 1 (LAPACK INFO)."""
    x = np.zeros((5, 1))
    y = np.ones(5)
    spec = _constant_spec()
    assert oracle.loglik(x, y, None, spec, [1.0], noise=-0.5)[1] == 2
    assert oracle.loglik(x, y, None, spec, [1.0], noise=-1.0)[1] == 1
    assert oracle.loglik(x, y, None, spec, [1.0], noise=-3.0)[1] == 1
    assert oracle.loglik(x, y, None, spec, [np.nan], noise=1.0)[1] == 1
    _, info = oracle.potrf(np.diag([1.0, 2.0, 0.0, 4.0]))
    assert info == 3


def test_potrf_potrs_against_numpy():
    rng = np.random.default_rng(9)
    a = rng.normal(size=(90, 90))
    k = a @ a.T + 90 * np.eye(90)
    l, info = oracle.potrf(k)
    assert info == 0
    np.testing.assert_allclose(l, np.linalg.cholesky(k), rtol=1e-12, atol=1e-12)
    b = rng.normal(size=90)
    np.testing.assert_allclose(oracle.potrs(l, b), np.linalg.solve(k, b), rtol=1e-11)
    np.testing.assert_allclose(oracle.logdet(l), np.linalg.slogdet(k)[1], rtol=1e-13)


def test_kernel_special_values():
    x = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0 + 2 * np.pi], [0.0, 0.0, 1.0 + np.pi]])
    per = oracle.Spec([LEAF_PERIODIC], [1], [3])
    th = np.array([0.4])
    k = oracle.kernel_matrix(x, per, th)
    lp = np.exp(-2 * th[0])
    assert abs(k[1, 0] - 1.0) < 1e-15          # a full period apart: sin(pi)^2 ~ 0
    np.testing.assert_allclose(k[2, 0], np.exp(-2 * lp), rtol=1e-14)   # half period: sin(pi/2) = 1
    assert np.all(np.diag(k) == 1.0)
    rbf = oracle.Spec([LEAF_RBF], [2], [1, 3])
    xr = np.array([[0.0, 5.0, 0.0], [1.0, -5.0, 2.0]])
    k = oracle.kernel_matrix(xr, rbf, [np.log(3.0), np.log(0.5)])
    np.testing.assert_allclose(k[1, 0], np.exp(-0.5 * (3.0 * 1.0 + 0.5 * 4.0)), rtol=1e-15)
    k = oracle.kernel_matrix(xr, rbf, [-800.0, -800.0])  # l -> 0
    assert k[1, 0] == 1.0
    cyc = oracle.Spec([LEAF_RBF_CYCLIC], [3], [1, 2, 3])
    xc = np.array([[0.0, 0.0, -3.0], [0.0, 0.0, 3.0]])   # 6 apart, wraps to 6 - 2 pi
    k = oracle.kernel_matrix(xc, cyc, [0.0, 0.0, np.log(2.0)])
    np.testing.assert_allclose(k[1, 0], np.exp(-0.5 * 2.0 * (6 - 2 * np.pi) ** 2), rtol=1e-13)
    lin = oracle.Spec([LEAF_LINEAR], [2], [1, 2])
    k = oracle.kernel_matrix(xr, lin, [0.3, -0.2])
    np.testing.assert_allclose(k[1, 0], np.exp(0.3) + np.exp(-0.2), rtol=1e-15)


def test_parameter_order_and_composition():
    """theta is stacked k1 then k2 (kernel_module.f90:227-245); sum / product
    combine the children elementwise."""
    rng = np.random.default_rng(4)
    x = rng.uniform(-2, 2, (7, 3))
    s_sum = oracle.Spec([LEAF_RBF, LEAF_PERIODIC, OP_SUM], [2, 1, 0], [1, 2, 3])
    s_prod = oracle.Spec([LEAF_RBF, LEAF_PERIODIC, OP_PROD], [2, 1, 0], [1, 2, 3])
    k1 = oracle.kernel_matrix(x, oracle.Spec([LEAF_RBF], [2], [1, 2]), [0.1, 0.2])
    k2 = oracle.kernel_matrix(x, oracle.Spec([LEAF_PERIODIC], [1], [3]), [0.3])
    th = [0.1, 0.2, 0.3]
    np.testing.assert_allclose(oracle.kernel_matrix(x, s_sum, th), k1 + k2, rtol=1e-15)
    np.testing.assert_allclose(oracle.kernel_matrix(x, s_prod, th), k1 * k2, rtol=1e-15)
    assert oracle.nparams(s_sum) == 3 and oracle.nparams(s_sum, True) == 4


def test_bad_specs_rejected():
    assert oracle.nparams(oracle.Spec([OP_SUM], [0], [])) == -1
    assert oracle.nparams(oracle.Spec([LEAF_RBF, LEAF_RBF], [1, 1], [1, 1])) == -1
    with pytest.raises(ValueError):
        oracle.kernel_matrix(np.zeros((3, 2)), oracle.Spec([LEAF_RBF], [1], [3]), [0.0])


def test_brute_force_long_double():
    """Tiny case recomputed in 80-bit extended precision."""
    rng = np.random.default_rng(11)
    n = 6
    x = rng.uniform(1.8, 2.9, (n, 3))
    y = rng.normal(size=n)
    spec = oracle.Spec([LEAF_RBF, LEAF_PERIODIC, OP_SUM], [2, 1, 0], [1, 2, 3])
    th = np.array([0.5, 1.0, 0.2])
    ld = np.longdouble
    xl = x.astype(ld)
    k = np.zeros((n, n), dtype=ld)
    for i in range(n):
        for j in range(n):
            d = xl[i] - xl[j]
            e1 = -ld(0.5) * (np.exp(ld(th[0])) * d[0] ** 2 + np.exp(ld(th[1])) * d[1] ** 2)
            e2 = -ld(2) * np.exp(-ld(2) * ld(th[2])) * np.sin(d[2] / ld(2)) ** 2
            k[i, j] = np.exp(e1) + np.exp(e2)
    k += ld(1e-3) * np.eye(n, dtype=ld)
    l = np.zeros_like(k)
    for j in range(n):
        l[j, j] = np.sqrt(k[j, j] - np.sum(l[j, :j] ** 2))
        for i in range(j + 1, n):
            l[i, j] = (k[i, j] - np.sum(l[i, :j] * l[j, :j])) / l[j, j]
    yl = y.astype(ld)
    z = np.zeros(n, dtype=ld)
    for i in range(n):
        z[i] = (yl[i] - np.sum(l[i, :i] * z[:i])) / l[i, i]
    ll_ref = -ld(0.5) * np.sum(z * z) - np.sum(np.log(np.diag(l))) - ld(0.5) * n * np.log(ld(2) * ld(np.pi))
    ll, info, _, _ = oracle.loglik(x, y, None, spec, th, noise=1e-3)
    assert info == 0
    assert abs(ll - float(ll_ref)) <= 1e-12 * abs(float(ll_ref))


def test_gradient_finite_differences():
    rng = np.random.default_rng(21)
    n = 60
    x = np.column_stack([rng.uniform(1.8, 2.9, n), rng.uniform(1.6, 2.2, n), rng.uniform(-np.pi, np.pi, n)])
    y = np.sin(2 * x[:, 0]) + 0.3 * np.cos(x[:, 2]) + 0.01 * rng.normal(size=n)
    spec = oracle.Spec([LEAF_RBF, LEAF_PERIODIC, OP_SUM, LEAF_CONSTANT, OP_PROD], [2, 1, 0, 0, 0], [1, 2, 3])
    m = np.full(n, y.mean())
    for lik in ("marginal", "concentrated"):
        th = np.array([0.7, 0.3, 0.1, 1.3, 1e-2])
        ll, g, info = oracle.loglik_grad(x, y, m, spec, th, optimise_noise=True, likelihood=lik)
        assert info == 0
        fd = np.empty_like(th)
        for q in range(len(th)):
            h = 1e-6 * max(abs(th[q]), 1e-2)
            e = np.zeros_like(th)
            e[q] = h
            fp = oracle.loglik(x, y, m, spec, th + e, optimise_noise=True, likelihood=lik)[0]
            fm = oracle.loglik(x, y, m, spec, th - e, optimise_noise=True, likelihood=lik)[0]
            fd[q] = (fp - fm) / (2 * h)
        np.testing.assert_allclose(g, fd, rtol=2e-5, atol=1e-6 * np.abs(fd).max())


def test_baseline_matches_oracle():
    """The CPU baseline (reference host algorithm + LAPACK) agrees with the oracle."""
    rng = np.random.default_rng(8)
    n = 120
    x = np.column_stack([rng.uniform(1.8, 2.9, n), rng.uniform(1.6, 2.2, n), rng.uniform(-np.pi, np.pi, n)])
    y = np.cos(x[:, 0] * 2) + 0.1 * np.sin(x[:, 2])
    spec = oracle.Spec([LEAF_RBF, LEAF_PERIODIC, OP_SUM], [2, 1, 0], [1, 2, 3])
    m = np.full(n, y.mean())
    th = rng.uniform(0.0, 1.0, (4, 3))
    ll, info = oracle.swarm_loglik(x, y, m, spec, th, noise=1e-6)
    for path in (oracle.scipy_openblas_path(), None):
        bl = oracle.Baseline(x, y, m, spec, noise=1e-6, lapack_path=path)
        bll, binfo = bl.swarm(th, threads=2)
        assert np.array_equal(info, binfo)
        np.testing.assert_allclose(bll, ll, rtol=1e-9)
        bl.close()
