#!/usr/bin/env python3
"""Writes tests/golden/golden_lapack.npz: log-likelihood golden vectors.

The vectors follow the reference's own host algorithm, but computed with the
LAPACK/BLAS routines it calls (scipy's bundled OpenBLAS/LAPACK):

* per-feature pair tensors D = (xi - xj)^2 and P = sin(sqrt(D)/2)^2
  (src/kernels/distance_matrix.f90:64-142), cyclic wrap on features 3, 6, ...
  (:113-125);
* R = exp(dgemv('t', D_active, -1/2 * l)) for rbf and rbf-cyclic,
  exp(dgemv('t', P_active, -2 * l)) for periodic
  (src/kernels/stationary_kernel.f90:101-122, periodic_kernel.f90:181-206),
  sum / product of children (kernel_module.f90:309-375), constant leaves
  (constant.f90:62-90), K += nugget I (gpr.f90:185-189);
* LAPACK dpotrf('L') / dpotrs('L') (src/utils/linalg.f90:73-168) and the
  marginal / concentrated likelihood (src/model/gpr/likelihood.f90:161-253).

The cases are small (N <= 48) so the fixture stays a few tens of KB.
Run from the repository root:  python tests/golden/make_golden.py
"""
import json
from pathlib import Path

import numpy as np
from scipy.linalg import blas, lapack

OUT = Path(__file__).resolve().parent / "golden_lapack.npz"
PI = 4.0 * np.arctan(1.0)


def tensors(x):
    n, d = x.shape
    diff = x[:, None, :] - x[None, :, :]                       # (i, j, f)
    dd = diff * diff
    per = np.sin(np.sqrt(dd) / 2.0) ** 2
    cyc = diff.copy()
    idx = np.arange(2, d, 3)
    cyc[:, :, idx] = np.sign(cyc[:, :, idx]) * (np.fmod(np.abs(cyc[:, :, idx]) + PI, 2 * PI) - PI)
    return dd, per, cyc * cyc


def leaf_matrix(kind, dims, th, x, T):
    n = x.shape[0]
    dd, per, cyc = T
    if kind in ("rbf", "periodic", "rbf-cyclic"):
        src = {"rbf": dd, "periodic": per, "rbf-cyclic": cyc}[kind]
        a = np.ascontiguousarray(src[:, :, [f - 1 for f in dims]].reshape(n * n, len(dims)).T)
        if kind == "periodic":
            l = (1.0 / np.exp(th)) ** 2
            alpha = -2.0
        else:
            l = np.exp(th)
            alpha = -0.5
        r = blas.dgemv(alpha, np.asfortranarray(a), l, trans=1)
        return np.exp(r).reshape(n, n)
    if kind == "linear":
        return np.full((n, n), np.sum(1.0 / np.exp(-th)))
    if kind == "constant":
        return np.full((n, n), th[0])
    raise ValueError(kind)


def build_k(x, program, theta, noise):
    """program: postfix list of ('leaf', kind, dims) / ('op', '+'|'*')."""
    T = tensors(x)
    st, off = [], 0
    for node in program:
        if node[0] == "op":
            b, a = st.pop(), st.pop()
            st.append(a + b if node[1] == "+" else a * b)
        else:
            kind, dims = node[1], node[2]
            npar = 1 if kind == "constant" else len(dims)
            st.append(leaf_matrix(kind, dims, theta[off:off + npar], x, T))
            off += npar
    k = st[0] + noise * np.eye(x.shape[0])
    return k, off


def lapack_ll(k, y, m, likelihood):
    n = k.shape[0]
    c, info = lapack.dpotrf(k, lower=1)
    if info != 0:
        return np.nan, int(info), np.zeros(n), np.full(4, np.nan)
    logdet = 2.0 * np.sum(np.log(np.diag(c)))
    if likelihood == "marginal":
        r = y - m
        a, _ = lapack.dpotrs(c, r, lower=1)
        ll = -0.5 * r @ a - 0.5 * logdet - 0.5 * n * np.log(2 * PI)
        return ll, 0, a, np.array([ll, logdet, 0.0, 0.0])
    ry, _ = lapack.dpotrs(c, y, lower=1)
    r1, _ = lapack.dpotrs(c, np.ones(n), lower=1)
    mu = np.sum(ry) / np.sum(r1)
    r = y - mu
    a, _ = lapack.dpotrs(c, r, lower=1)
    s2 = r @ a / n
    ll = -0.5 * n * np.log(s2) - 0.5 * logdet
    return ll, 0, a, np.array([ll, logdet, mu, s2])


def cases():
    rng = np.random.default_rng(2024)
    out = []
    n = 40
    x = np.column_stack([rng.normal(1.81, 0.05, n), rng.normal(1.81, 0.05, n), rng.normal(1.82, 0.08, n)])
    y = np.sin(3 * x[:, 0]) + 0.5 * np.cos(2 * x[:, 2]) + 0.01 * rng.normal(size=n)
    prog = [("leaf", "rbf", [1, 2]), ("leaf", "periodic", [3]), ("op", "+")]
    for lik in ("marginal", "concentrated"):
        for t in range(3):
            th = rng.uniform(1e-3, 1.0, 3) + np.array([2.0, 2.0, 0.0])
            out.append(dict(name=f"water-{lik}-{t}", x=x, y=y, mean="constant", program=prog, theta=th,
                            noise=1e-6, optimise_noise=False, likelihood=lik))
    d, n = 6, 48
    x6 = np.column_stack([rng.uniform(1.8, 2.9, n), rng.uniform(1.8, 2.9, n), rng.uniform(1.6, 2.2, n),
                          rng.uniform(1.8, 2.9, n), rng.uniform(1.6, 2.2, n), rng.uniform(-PI, PI, n)])
    y6 = np.exp(-np.sum((x6 - 2.0) ** 2, axis=1)) + 0.02 * np.cos(x6[:, 5])
    prog2 = [("leaf", "rbf", [1, 2]), ("leaf", "rbf", [3, 4]), ("op", "+"), ("leaf", "periodic", [6]), ("op", "*")]
    for t in range(2):
        th = rng.uniform(0.0, 1.0, 5)
        out.append(dict(name=f"nested-{t}", x=x6, y=y6, mean="linear", program=prog2, theta=th, noise=1e-7,
                        optimise_noise=False, likelihood="marginal"))
    prog3 = [("leaf", "rbf-cyclic", [1, 2, 3, 4, 5, 6])]
    th = rng.uniform(0.0, 1.0, 6)
    out.append(dict(name="cyclic", x=x6, y=y6, mean="quadratic", program=prog3, theta=th, noise=1e-6,
                    optimise_noise=False, likelihood="marginal"))
    prog4 = [("leaf", "constant", []), ("leaf", "rbf", [1, 2, 4]), ("op", "*"), ("leaf", "linear", [3]), ("op", "+")]
    th = np.array([1.7, 0.2, 0.4, 0.9, -0.3])
    out.append(dict(name="scaled-linear", x=x6, y=y6, mean="zero", program=prog4, theta=th, noise=1e-5,
                    optimise_noise=False, likelihood="concentrated"))
    prog5 = [("leaf", "rbf", [1, 2, 3, 4, 5]), ("leaf", "periodic", [6]), ("op", "+")]
    for t, nug in enumerate([1e-4, -3.0, -1e-9]):
        th = np.concatenate([rng.uniform(0.0, 1.0, 6), [nug]])
        out.append(dict(name=f"noise-{t}", x=x6, y=y6, mean="constant", program=prog5, theta=th, noise=1e-10,
                        optimise_noise=True, likelihood="marginal"))
    return out


def mean_value(kind, x, y):
    n = len(y)
    if kind == "zero":
        return np.zeros(n)
    if kind == "constant":
        return np.full(n, np.sum(y) / n)
    imin = int(np.argmin(y))
    a = x - x[imin]
    if kind == "quadratic":
        a = a * a
    c, info = lapack.dpotrf(a.T @ a, lower=1)
    beta, _ = lapack.dpotrs(c, a.T @ (y - y[imin]), lower=1)
    return a @ beta + y[imin]


def main():
    arrays, meta = {}, []
    for k, c in enumerate(cases()):
        m = mean_value(c["mean"], c["x"], c["y"])
        noise = c["theta"][-1] if c["optimise_noise"] else c["noise"]
        K, npar = build_k(c["x"], c["program"], c["theta"], noise)
        ll, info, alpha, out4 = lapack_ll(K, c["y"], m, c["likelihood"])
        cond = float(np.linalg.cond(K)) if info == 0 else float("inf")
        for key in ("x", "y", "theta"):
            arrays[f"{k}_{key}"] = c[key]
        arrays[f"{k}_mean"] = m
        arrays[f"{k}_alpha"] = alpha
        arrays[f"{k}_out4"] = out4
        meta.append(dict(name=c["name"], mean=c["mean"], likelihood=c["likelihood"], noise=c["noise"],
                         optimise_noise=c["optimise_noise"], info=info, ll=ll, cond=cond,
                         program=[list(p) for p in c["program"]]))
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} with {len(meta)} cases")


if __name__ == "__main__":
    main()
