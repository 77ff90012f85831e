"""Swarm log-likelihood operator over the gfx950 library.

Host mirror of the reference's GaussianProcessRegressor as the PSO sees it
(src/model/gpr/gpr.f90): construct once per training set, then evaluate the
log-likelihood of whole swarms of theta vectors.  Every evaluation runs in
libferebus_hip.so; nothing here computes a likelihood.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence, Tuple, Union

import numpy as np

from . import _abi
from .kernel_expr import KernelConfig, KernelProgram, compile_kernel
from .mean import MeanFit, fit_mean

LIKELIHOODS = {"marginal": 0, "concentrated": 1}
# lock-step factorisation per 128-column panel: "update" = trailing update +
# first diagonal block, "solve_factor" = panel solve + second diagonal block,
# "update_solve" = second block's update + solve
PROFILE_KINDS = {"build": 0, "update": 1, "solve_factor": 2, "update_solve": 3, "likelihood": 4, "swarm": 5,
                 "factor": 6, "tune_evals": 7, "factor_span": 8, "dispatch_fallbacks": 9}


class SwarmLikelihood:
    """Log-likelihood of every particle of a swarm for one training set.

    x        (N, d) training inputs (ALF features)
    y        (N,) training outputs
    kernel   kernel expression (``"k1+k2"``) with ``kernels`` configs, or a
             compiled :class:`KernelProgram`
    mean     mean kind (``zero|constant|linear|quadratic``), a fitted
             :class:`MeanFit`, or an explicit (N,) mean vector
    noise    nugget (gpr.f90:185-189); with ``optimise_noise`` the nugget is
             the last theta entry instead (gpr.f90:191-202)
    """

    def __init__(self, x: np.ndarray, y: np.ndarray, kernel: Union[str, KernelProgram],
                 kernels: Optional[Dict[str, KernelConfig]] = None,
                 mean: Union[str, MeanFit, np.ndarray, None] = "constant",
                 noise: float = 1e-10, optimise_noise: bool = False,
                 likelihood: str = "marginal", max_swarm: int = 64):
        self._lib = _abi.load()
        self._h = ctypes.c_void_p()
        x = np.asarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.ndim != 2 or y.shape != (x.shape[0],):
            raise ValueError("x must be (N, d) and y (N,)")
        self.n, self.d = x.shape
        if isinstance(kernel, str):
            kernel = compile_kernel(kernel, kernels or {}, self.d)
        self.program = kernel
        if isinstance(mean, str):
            mean = fit_mean(mean, x, y)
        if isinstance(mean, MeanFit):
            self.mean_fit = mean
            mvec = mean.value
        elif mean is None:
            self.mean_fit = None
            mvec = None
        else:
            self.mean_fit = None
            mvec = np.asarray(mean, dtype=np.float64)
        if likelihood not in LIKELIHOODS:
            raise ValueError(f"Unknown likelihood function: {likelihood}")
        self.likelihood = likelihood
        self.noise = float(noise)
        self.optimise_noise = bool(optimise_noise)
        self.max_swarm = int(max_swarm)
        self._x = np.asfortranarray(x)
        self._y = y
        self._m = None if mvec is None else np.ascontiguousarray(mvec, dtype=np.float64)
        nt = np.ascontiguousarray(kernel.node_type, dtype=np.int32)
        nd = np.ascontiguousarray(kernel.node_ndim, dtype=np.int32)
        ad = np.ascontiguousarray(kernel.active_dims if kernel.active_dims else [0], dtype=np.int32)
        self._spec_arrays = (nt, nd, ad)
        spec = _abi.KernelSpec(len(nt), _abi.iptr(nt), _abi.iptr(nd), _abi.iptr(ad))
        rc = self._lib.fb_create(ctypes.byref(self._h), self.n, self.d, _abi.dptr(self._x),
                                 _abi.dptr(self._y), _abi.dptr(self._m) if self._m is not None else None,
                                 ctypes.byref(spec), self.noise, int(self.optimise_noise),
                                 LIKELIHOODS[likelihood], self.max_swarm)
        _abi.check(rc, "fb_create")
        p = ctypes.c_int()
        _abi.check(self._lib.fb_nparams(self._h, ctypes.byref(p)), "fb_nparams")
        self.nparams = p.value

    # -- evaluation -------------------------------------------------------
    def _theta(self, theta) -> np.ndarray:
        t = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, dtype=np.float64)))
        if t.shape[1] != self.nparams:
            raise ValueError(f"theta must have {self.nparams} entries per particle, got {t.shape[1]}")
        return t

    def swarm_loglik(self, theta) -> Tuple[np.ndarray, np.ndarray]:
        """Batched path: (ll[S], info[S]) for theta of shape (S, P)."""
        t = self._theta(theta)
        s = t.shape[0]
        ll = np.empty(s)
        info = np.empty(s, dtype=np.int32)
        _abi.check(self._lib.fb_swarm_loglik(self._h, s, _abi.dptr(t), _abi.dptr(ll), _abi.iptr(info)),
                   "fb_swarm_loglik")
        return ll, info

    def particle_loglik(self, theta) -> Tuple[np.ndarray, np.ndarray]:
        """Plain path: one launch sequence per particle."""
        t = self._theta(theta)
        s = t.shape[0]
        ll = np.empty(s)
        info = np.empty(s, dtype=np.int32)
        _abi.check(self._lib.fb_particle_loglik(self._h, s, _abi.dptr(t), _abi.dptr(ll), _abi.iptr(info)),
                   "fb_particle_loglik")
        return ll, info

    def swarm_loglik_async(self, d_theta: int, d_ll: int, d_info: int, s: int) -> None:
        """Device-pointer variant: enqueue without synchronising."""
        _abi.check(self._lib.fb_swarm_loglik_async(self._h, s, ctypes.c_void_p(d_theta),
                                                   ctypes.c_void_p(d_ll), ctypes.c_void_p(d_info)),
                   "fb_swarm_loglik_async")

    def alpha(self, theta):
        """Weights alpha and (ll, logdet, mu_hat, sigma2_hat) for one theta."""
        t = self._theta(theta)[0].copy()
        alpha = np.empty(self.n)
        out4 = np.empty(4)
        info = ctypes.c_int()
        _abi.check(self._lib.fb_get_alpha(self._h, _abi.dptr(t), _abi.dptr(alpha), _abi.dptr(out4),
                                          ctypes.byref(info)), "fb_get_alpha")
        return alpha, out4, info.value

    def matrix(self, theta, stage: str = "K"):
        """The device's K(theta) (stage "K", nugget included) or its Cholesky
        factor (stage "L") for one theta: ((N, N) lower triangle, INFO)."""
        t = self._theta(theta)[0].copy()
        out = np.zeros((self.n, self.n), order="F")
        info = ctypes.c_int()
        _abi.check(self._lib.fb_debug_matrix(self._h, _abi.dptr(t), {"K": 0, "L": 1}[stage],
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             ctypes.byref(info)), "fb_debug_matrix")
        return out, info.value

    def swarm_grad(self, theta):
        """(ll[S], grad[S, P], info[S]) of the adjoint gradient path."""
        t = self._theta(theta)
        s = t.shape[0]
        ll = np.empty(s)
        grad = np.empty((s, self.nparams))
        info = np.empty(s, dtype=np.int32)
        _abi.check(self._lib.fb_swarm_loglik_grad(self._h, s, _abi.dptr(t), _abi.dptr(ll),
                                                  _abi.dptr(grad), _abi.iptr(info)),
                   "fb_swarm_loglik_grad")
        return ll, grad, info

    def cost_function(self, theta) -> Tuple[float, bool]:
        """negative_log_likelihood (gpr.f90:258-289): (-ll, success)."""
        ll, info = self.swarm_loglik(np.atleast_2d(theta))
        return -float(ll[0]), bool(info[0] == 0)

    # -- plumbing ---------------------------------------------------------
    def set_stream(self, stream_handle: Optional[int]) -> None:
        _abi.check(self._lib.fb_set_stream(self._h, ctypes.c_void_p(stream_handle or 0)), "fb_set_stream")

    FACTOR_PATHS = {-1: "untuned", 0: "persistent", 1: "lockstep-kc32", 2: "lockstep-kc16",
                    3: "lockstep-kc16-32x128", 4: "lockstep-kc16-32x64", 5: "lockstep-kc16-per-panel"}

    def factor_path(self, s: int) -> str:
        """Factorisation path chosen for a swarm of s (see fb_factor_path)."""
        p = ctypes.c_int(-1)
        _abi.check(self._lib.fb_factor_path(self._h, int(s), ctypes.byref(p)), "fb_factor_path")
        return self.FACTOR_PATHS[p.value]

    def factor_groups(self, s: int) -> int:
        """Stream groups of the lock-step factorisation for a swarm of s (see
        fb_factor_groups)."""
        g = ctypes.c_int(1)
        _abi.check(self._lib.fb_factor_groups(self._h, int(s), ctypes.byref(g)), "fb_factor_groups")
        return g.value

    def profile(self, enable) -> None:
        """True / 1: an event pair around every launch; 2: around the phases
        only (build, factorisation span, likelihood); False / 0: off."""
        _abi.check(self._lib.fb_profile_enable(self._h, int(enable)), "fb_profile_enable")

    def profile_read(self, kind: str) -> Tuple[float, int]:
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _abi.check(self._lib.fb_profile_read(self._h, PROFILE_KINDS[kind], ctypes.byref(ms), ctypes.byref(n)),
                   "fb_profile_read")
        return ms.value, n.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.fb_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiSwarm:
    """The swarms of several training sets (one SwarmLikelihood each, any N)
    evaluated together (fb_multi_swarm_loglik): builds and likelihoods per
    training set, the factorisations of all of them on the schedule the
    first call timed fastest -- the sets one after another, one persistent
    launch, or the chained lock-step launches shared by every set and
    replayed as one hipGraph.  Results are bit-identical to each
    SwarmLikelihood.swarm_loglik."""

    def __init__(self, gps: Sequence[SwarmLikelihood]):
        self._lib = _abi.load()
        self.gps = list(gps)
        arr = (ctypes.c_void_p * len(self.gps))(*[g._h.value for g in self.gps])
        self._h = ctypes.c_void_p()
        self._last = np.zeros(len(self.gps), dtype=np.int32)
        _abi.check(self._lib.fb_multi_create(ctypes.byref(self._h), len(self.gps), arr), "fb_multi_create")

    def swarm_loglik(self, thetas):
        """thetas: one (S_g, P_g) array per training set (S_g may be 0).
        Returns ([ll_g], [info_g])."""
        G = len(self.gps)
        ts = [np.zeros((0, g.nparams)) if t is None else g._theta(t) if len(np.atleast_2d(t)) else
              np.zeros((0, g.nparams)) for g, t in zip(self.gps, thetas)]
        s = np.array([len(t) for t in ts], dtype=np.int32)
        self._last = s
        lls = [np.empty(len(t)) for t in ts]
        infos = [np.empty(len(t), dtype=np.int32) for t in ts]
        P_DBL, P_INT = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        th_p = (P_DBL * G)(*[_abi.dptr(t) if len(t) else None for t in ts])
        ll_p = (P_DBL * G)(*[_abi.dptr(v) if len(v) else None for v in lls])
        in_p = (P_INT * G)(*[_abi.iptr(v) if len(v) else None for v in infos])
        _abi.check(self._lib.fb_multi_swarm_loglik(self._h, _abi.iptr(s), th_p, ll_p, in_p),
                   "fb_multi_swarm_loglik")
        return lls, infos

    SCHEDULES = ("sequential", "persistent", "merged")

    def pso_step(self, steps):
        """One iteration of several device swarms (fb_multi_pso_step).
        steps: {g: (swarm, r1, r2)} with `swarm` a DeviceSwarm on training
        set g and r1, r2 of shape (S_g, P_g).  Returns ({g: cost},
        {g: INFO})."""
        G = len(self.gps)
        P_DBL, P_INT = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        ps = (ctypes.c_void_p * G)()
        r1p, r2p, cp = (P_DBL * G)(), (P_DBL * G)(), (P_DBL * G)()
        ip = (P_INT * G)()
        keep, costs, infos = [], {}, {}
        for g, (sw, r1, r2) in steps.items():
            if sw.gp is not self.gps[g]:
                raise ValueError(f"swarm {g} is not on training set {g}")
            r1, r2 = sw._rows(r1, sw.S), sw._rows(r2, sw.S)
            costs[g], infos[g] = np.empty(sw.S), np.empty(sw.S, dtype=np.int32)
            keep += [r1, r2]
            ps[g] = sw._p.value
            r1p[g], r2p[g] = _abi.dptr(r1), _abi.dptr(r2)
            cp[g], ip[g] = _abi.dptr(costs[g]), _abi.iptr(infos[g])
        self._last = np.array([steps[g][0].S if g in steps else 0 for g in range(G)], dtype=np.int32)
        _abi.check(self._lib.fb_multi_pso_step(self._h, ps, r1p, r2p, cp, ip), "fb_multi_pso_step")
        return costs, infos

    def schedule(self, sizes=None):
        """The schedule chosen for the swarm sizes of the last call (or
        `sizes`): "sequential", "persistent", "merged", or None before a call
        with those sizes."""
        s = np.array(self._last if sizes is None else sizes, dtype=np.int32)
        k = self._lib.fb_multi_schedule(self._h, _abi.iptr(s))
        return None if k < 0 else self.SCHEDULES[k]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.fb_multi_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
