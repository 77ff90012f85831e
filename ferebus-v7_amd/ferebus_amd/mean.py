"""Prior mean functions, evaluated once per training set on the host.

Mirror of src/mean/ of the reference: the mean does not depend on theta, so
GaussianProcessRegressor%init caches its value (src/model/gpr/gpr.f90:168-171)
and the device only ever sees y - m.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

MEAN_KINDS = ("zero", "constant", "linear", "quadratic")


@dataclass
class MeanFit:
    kind: str
    value: np.ndarray                      # m(x_i), length N
    constant: float = 0.0                  # constant mean value
    beta: Optional[np.ndarray] = None      # least-squares means
    xmin: Optional[np.ndarray] = None
    ymin: float = 0.0
    info: int = 0                          # dposv INFO of the normal equations
    extra: dict = field(default_factory=dict)


def _cholesky_solve(a: np.ndarray, b: np.ndarray):
    """dposv('L') of the normal equations (src/utils/linalg.f90:205-238).
    Returns (solution, INFO)."""
    n = a.shape[0]
    l = np.array(a, dtype=np.float64, copy=True)
    for j in range(n):
        ajj = l[j, j] - np.dot(l[j, :j], l[j, :j])
        if not ajj > 0.0:
            return np.array(b, dtype=np.float64, copy=True), j + 1
        ajj = np.sqrt(ajj)
        l[j, j] = ajj
        if j + 1 < n:
            l[j + 1:, j] = (l[j + 1:, j] - l[j + 1:, :j] @ l[j, :j]) / ajj
    z = np.array(b, dtype=np.float64, copy=True)
    for j in range(n):
        z[j] = (z[j] - np.dot(l[j, :j], z[:j])) / l[j, j]
    for j in range(n - 1, -1, -1):
        z[j] = (z[j] - np.dot(l[j + 1:, j], z[j + 1:])) / l[j, j]
    return z, 0


def fit_mean(kind: str, x: np.ndarray, y: np.ndarray) -> MeanFit:
    """Fit and evaluate a mean on the training inputs.

    zero       src/mean/zero.f90:47-52
    constant   src/mean/constant.f90:42-55 (mean(y), src/utils/arith.f90:44-48)
    linear     src/mean/least_squares.f90:54-93 + src/mean/linear.f90:41-51
    quadratic  src/mean/least_squares.f90:54-93 + src/mean/quadratic.f90:41-52
    The least-squares means fit y - y_min on x - x_min (x_min = the point of
    smallest y, first occurrence), without an intercept; the quadratic mean
    uses only the squared terms.
    """
    kind = kind.lower()
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = y.shape[0]
    if kind == "zero":
        return MeanFit("zero", np.zeros(n))
    if kind == "constant":
        c = float(np.sum(y) / n)
        return MeanFit("constant", np.full(n, c), constant=c)
    if kind not in ("linear", "quadratic"):
        raise ValueError(f"Unknown Mean Type: {kind}")
    imin = int(np.argmin(y))
    xmin = x[imin, :].copy()
    ymin = float(y[imin])
    a = x - xmin
    if kind == "quadratic":
        a = a * a
    beta, info = _cholesky_solve(a.T @ a, a.T @ (y - ymin))
    value = a @ beta + ymin
    return MeanFit(kind, value, beta=beta, xmin=xmin, ymin=ymin, info=info)
