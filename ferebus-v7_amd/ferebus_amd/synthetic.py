"""Seeded synthetic training sets shaped like the benchmark systems.

There is no network and no reference output data, so every workload is
generated here with numpy from a fixed seed.  ALF feature order follows the
reference (r1, r2, angle, then distance, angle, dihedral per further atom;
src/data/units.f90:55-64): dihedrals sit at 1-based features 6, 9, 12, ...
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from .kernel_expr import KernelConfig


@dataclass
class Workload:
    name: str
    x: np.ndarray                     # (N, d)
    y: np.ndarray                     # (N,)
    kernel: str
    kernels: Dict[str, KernelConfig]
    mean: str = "constant"
    likelihood: str = "marginal"
    noise: float = 1e-10
    optimise_noise: bool = False
    swarm: int = 50
    iterations: int = 100
    search_min: float = 0.0
    search_max: float = 1.0
    extra: dict = field(default_factory=dict)


def _feature_kinds(d: int) -> List[str]:
    """'dist' / 'angle' / 'dihedral' for ALF feature positions 1..d."""
    kinds = ["dist", "dist", "angle"]
    while len(kinds) < d:
        kinds += ["dist", "angle", "dihedral"]
    return kinds[:d]


def alf_features(rng: np.random.Generator, n: int, d: int) -> np.ndarray:
    """Distances U[1.8, 2.9] bohr, angles U[1.6, 2.2] rad, dihedrals U[-pi, pi]."""
    x = np.empty((n, d))
    for f, kind in enumerate(_feature_kinds(d)):
        if kind == "dist":
            x[:, f] = rng.uniform(1.8, 2.9, n)
        elif kind == "angle":
            x[:, f] = rng.uniform(1.6, 2.2, n)
        else:
            x[:, f] = rng.uniform(-np.pi, np.pi, n)
    return x


def bump_target(rng: np.random.Generator, x: np.ndarray, nbumps: int = 20) -> np.ndarray:
    """Sum of Gaussian bumps plus 0.005 cos(k phi), k = 1..3, over the dihedrals."""
    n, d = x.shape
    centres = alf_features(rng, nbumps, d)
    widths = rng.uniform(0.3, 1.0, nbumps)
    amps = rng.normal(0.0, 0.01, nbumps)
    y = np.zeros(n)
    for c, w, a in zip(centres, widths, amps):
        y += a * np.exp(-np.sum((x - c) ** 2, axis=1) / (2.0 * w * w))
    kinds = _feature_kinds(d)
    for f, kind in enumerate(kinds):
        if kind == "dihedral":
            for k in (1, 2, 3):
                y += 0.005 * np.cos(k * x[:, f])
    return y


def rbf_periodic_kernels(d: int, ard: bool = True) -> Dict[str, KernelConfig]:
    kinds = _feature_kinds(d)
    rbf_dims = [f + 1 for f, k in enumerate(kinds) if k != "dihedral"]
    per_dims = [f + 1 for f, k in enumerate(kinds) if k == "dihedral"]
    ks = {"k1": KernelConfig("k1", "rbf", rbf_dims)}
    if per_dims:
        ks["k2"] = KernelConfig("k2", "periodic", per_dims)
    return ks


def water(n: int = 500, seed: int = 0) -> Workload:
    """Water O1, IQA energy: r1, r2 ~ N(1.81, 0.05), angle ~ N(1.82, 0.08);
    y = two Morse terms + 0.1 (theta - 1.82)^2 + N(0, 1e-6)."""
    rng = np.random.default_rng(seed)
    x = np.column_stack([rng.normal(1.81, 0.05, n), rng.normal(1.81, 0.05, n),
                         rng.normal(1.82, 0.08, n)])
    de, a, re = 0.2, 1.2, 1.81
    morse = lambda r: de * (1.0 - np.exp(-a * (r - re))) ** 2
    y = morse(x[:, 0]) + morse(x[:, 1]) + 0.1 * (x[:, 2] - 1.82) ** 2 + rng.normal(0.0, 1e-6, n)
    ks = {"k1": KernelConfig("k1", "rbf", [1, 2]), "k2": KernelConfig("k2", "periodic", [3])}
    return Workload("water-O1", x, y, "k1+k2", ks, swarm=20, iterations=50,
                    search_min=1e-3, search_max=1.0)


def methanol(n: int = 2000, seed: int = 0) -> Workload:
    """Methanol-like 6-atom system: N=2000, d=12, rbf on 9 + periodic on 3."""
    rng = np.random.default_rng(seed)
    d = 12
    x = alf_features(rng, n, d)
    y = bump_target(rng, x)
    return Workload("methanol-like", x, y, "k1+k2", rbf_periodic_kernels(d), swarm=50, iterations=100)


def glycine_gps(seed: int = 2, n_gps: int = 10, nmin: int = 1000, nmax: int = 4000) -> List[Workload]:
    """10 independent GPs of a glycine-like system, d=24, N ~ U{1000..4000}."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, n_gps)
    out = []
    for g, n in enumerate(sizes):
        x = alf_features(rng, int(n), 24)
        y = bump_target(rng, x)
        out.append(Workload(f"glycine-like-{g}", x, y, "k1+k2", rbf_periodic_kernels(24), swarm=40))
    return out


def nma(n: int = 3000, seed: int = 3) -> Workload:
    """NMA-like: d=30, ARD rbf on all 30 + periodic on the 9 dihedrals + noise."""
    rng = np.random.default_rng(seed)
    d = 30
    x = alf_features(rng, n, d)
    y = bump_target(rng, x)
    per = [f + 1 for f, k in enumerate(_feature_kinds(d)) if k == "dihedral"]
    ks = {"k1": KernelConfig("k1", "rbf", list(range(1, d + 1))), "k2": KernelConfig("k2", "periodic", per)}
    return Workload("nma-like", x, y, "k1+k2", ks, optimise_noise=True, swarm=32,
                    extra={"noise_search": (1e-12, 1e-4)})


def ibuprofen(n: int = 6000, seed: int = 4) -> Workload:
    """Ibuprofen-like: d=93, linear mean, concentrated likelihood."""
    rng = np.random.default_rng(seed)
    d = 93
    x = alf_features(rng, n, d)
    y = bump_target(rng, x)
    return Workload("ibuprofen-like", x, y, "k1+k2", rbf_periodic_kernels(d), mean="linear",
                    likelihood="concentrated", swarm=64, iterations=200)
