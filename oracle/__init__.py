"""CPU FP64 oracle of FEREBUS-v7's swarm log-likelihood path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The product
(libferebus_hip.so and the ferebus_amd package) never imports it.

Parity pins (see tests/test_oracle.py): scipy's LAPACK dpotrf / dpotrs golden
vectors (tests/golden/), closed forms (K = cI, K = a 11' + bI, periodic
kernel at a full period), brute force on tiny inputs, and the reference's
example training sets.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_P_DBL = ctypes.POINTER(ctypes.c_double)
_P_INT = ctypes.POINTER(ctypes.c_int)


class _Spec(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int), ("node_type", _P_INT),
                ("node_ndim", _P_INT), ("active_dims", _P_INT)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = ctypes.CDLL(str(LIB_PATH))
        sp = ctypes.POINTER(_Spec)
        lib.or_nparams.argtypes = [sp, ctypes.c_int]
        lib.or_kernel_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _P_DBL, sp, _P_DBL, _P_DBL, ctypes.c_int]
        lib.or_potrf.argtypes = [ctypes.c_int, _P_DBL, ctypes.c_int]
        lib.or_potrs.argtypes = [ctypes.c_int, _P_DBL, ctypes.c_int, _P_DBL]
        lib.or_potrs.restype = None
        lib.or_logdet.argtypes = [ctypes.c_int, _P_DBL, ctypes.c_int]
        lib.or_logdet.restype = ctypes.c_double
        lib.or_loglik.argtypes = [ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL, sp, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL]
        lib.or_swarm_loglik.argtypes = [ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL, sp,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P_DBL,
                                        _P_DBL, _P_INT, ctypes.c_int]
        lib.or_loglik_grad.argtypes = [ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL, sp, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL]
        lib.or_kernel_matrix_ext.argtypes = [ctypes.c_int, ctypes.c_int, _P_DBL, sp, _P_DBL, _P_DBL, _P_DBL,
                                             ctypes.c_int]
        lib.or_loglik_ext.argtypes = [ctypes.c_int, _P_DBL, ctypes.c_int, _P_DBL, _P_DBL, ctypes.c_int, _P_DBL,
                                      _P_DBL]
        lib.or_mean.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL, _P_DBL,
                                _P_DBL, _P_DBL]
        lib.or_baseline_create.argtypes = [ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL, sp,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                           ctypes.c_char_p]
        lib.or_baseline_create.restype = ctypes.c_void_p
        lib.or_baseline_lapack.argtypes = [ctypes.c_void_p]
        lib.or_baseline_swarm.argtypes = [ctypes.c_void_p, ctypes.c_int, _P_DBL, _P_DBL, _P_INT, ctypes.c_int]
        lib.or_baseline_destroy.argtypes = [ctypes.c_void_p]
        lib.or_baseline_destroy.restype = None
        _lib = lib
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_P_DBL)


class Spec:
    """Kernel program in fb_kernel_spec layout (node_type, node_ndim, active_dims)."""

    def __init__(self, node_type, node_ndim, active_dims):
        self.nt = np.ascontiguousarray(node_type, dtype=np.int32)
        self.nd = np.ascontiguousarray(node_ndim, dtype=np.int32)
        self.ad = np.ascontiguousarray(active_dims if len(active_dims) else [0], dtype=np.int32)
        self.c = _Spec(len(self.nt), self.nt.ctypes.data_as(_P_INT), self.nd.ctypes.data_as(_P_INT),
                       self.ad.ctypes.data_as(_P_INT))

    @classmethod
    def from_program(cls, prog):
        return cls(prog.node_type, prog.node_ndim, prog.active_dims)

    def ref(self):
        return ctypes.byref(self.c)


LIK = {"marginal": 0, "concentrated": 1}


def nparams(spec: Spec, optimise_noise=False) -> int:
    return load().or_nparams(spec.ref(), int(optimise_noise))


def kernel_matrix(x, spec: Spec, theta) -> np.ndarray:
    """Lower triangle (and diagonal) of K without nugget; upper part zero."""
    x = np.asfortranarray(x, dtype=np.float64)
    n, d = x.shape
    k = np.zeros((n, n), order="F")
    th = np.ascontiguousarray(theta, dtype=np.float64)
    rc = load().or_kernel_matrix(n, d, _d(x), spec.ref(), _d(th), _d(k), n)
    if rc:
        raise ValueError("bad kernel spec")
    return k


def potrf(a) -> tuple:
    """(L, INFO) of the LAPACK-dpotf2 restatement; a is read as its lower part."""
    l = np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
    n = l.shape[0]
    info = load().or_potrf(n, _d(l), n)
    return np.tril(l), info


def potrs(l, b) -> np.ndarray:
    l = np.asfortranarray(l, dtype=np.float64)
    x = np.array(b, dtype=np.float64, copy=True)
    load().or_potrs(l.shape[0], _d(l), l.shape[0], _d(x))
    return x


def logdet(l) -> float:
    l = np.asfortranarray(l, dtype=np.float64)
    return load().or_logdet(l.shape[0], _d(l), l.shape[0])


def loglik(x, y, mean, spec: Spec, theta, noise=1e-10, optimise_noise=False, likelihood="marginal"):
    """(ll, info, alpha, (ll, logdet, mu_hat, sigma2_hat)) of one theta."""
    x = np.asfortranarray(x, dtype=np.float64)
    n, d = x.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = None if mean is None else np.ascontiguousarray(mean, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    alpha = np.zeros(n)
    out4 = np.full(4, np.nan)
    info = load().or_loglik(n, d, _d(x), _d(y), _d(m), spec.ref(), noise, int(optimise_noise),
                            LIK[likelihood], _d(th), _d(alpha), _d(out4))
    if info < 0:
        raise ValueError("bad kernel spec")
    return out4[0], info, alpha, out4


def swarm_loglik(x, y, mean, spec: Spec, theta, noise=1e-10, optimise_noise=False,
                 likelihood="marginal", threads=0):
    x = np.asfortranarray(x, dtype=np.float64)
    n, d = x.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = None if mean is None else np.ascontiguousarray(mean, dtype=np.float64)
    th = np.ascontiguousarray(np.atleast_2d(theta), dtype=np.float64)
    s = th.shape[0]
    ll = np.empty(s)
    info = np.empty(s, dtype=np.int32)
    rc = load().or_swarm_loglik(n, d, _d(x), _d(y), _d(m), spec.ref(), noise, int(optimise_noise),
                                LIK[likelihood], s, _d(th), _d(ll), info.ctypes.data_as(_P_INT), threads)
    if rc:
        raise ValueError("bad kernel spec")
    return ll, info


def loglik_grad(x, y, mean, spec: Spec, theta, noise=1e-10, optimise_noise=False, likelihood="marginal"):
    """(ll, grad, info) with the dense reference gradient."""
    x = np.asfortranarray(x, dtype=np.float64)
    n, d = x.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = None if mean is None else np.ascontiguousarray(mean, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    p = nparams(spec, optimise_noise)
    grad = np.zeros(p)
    ll = ctypes.c_double(np.nan)
    info = load().or_loglik_grad(n, d, _d(x), _d(y), _d(m), spec.ref(), noise, int(optimise_noise),
                                 LIK[likelihood], _d(th), ctypes.byref(ll), _d(grad))
    return ll.value, grad, info


def kernel_matrix_ext(x, spec: Spec, theta):
    """The kernel matrix (lower triangle, no nugget) evaluated in long double:
    (k_hi, k_lo), the nearest FP64 values and their remainders."""
    x = np.asfortranarray(x, dtype=np.float64)
    n, d = x.shape
    hi = np.zeros((n, n), order="F")
    lo = np.zeros((n, n), order="F")
    th = np.ascontiguousarray(theta, dtype=np.float64)
    if load().or_kernel_matrix_ext(n, d, _d(x), spec.ref(), _d(th), _d(hi), _d(lo), n):
        raise ValueError("bad kernel spec")
    return hi, lo


def ulp_error(a, k_hi, k_lo):
    """Signed error of the FP64 matrix a against the extended truth
    (k_hi + k_lo), in units of the spacing of FP64 numbers at k_hi."""
    LD = np.longdouble
    err = (np.asarray(a, dtype=LD) - np.asarray(k_hi, dtype=LD)) - np.asarray(k_lo, dtype=LD)
    sp = np.spacing(np.abs(np.asarray(k_hi, dtype=np.float64)))
    return np.asarray(err / np.asarray(sp, dtype=LD), dtype=np.float64)


def loglik_ext(k, y, mean=None, likelihood="marginal", factor=False):
    """Extended-precision (80-bit long double) likelihood of the given FP64
    matrix k (lower triangle read, nugget included): (ll, info, out4) or,
    with factor=True, (ll, info, out4, L rounded to FP64)."""
    k = np.asfortranarray(k, dtype=np.float64)
    n = k.shape[0]
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = None if mean is None else np.ascontiguousarray(np.broadcast_to(mean, (n,)), dtype=np.float64)
    out4 = np.full(4, np.nan)
    lf = np.zeros((n, n), order="F") if factor else None
    info = load().or_loglik_ext(n, _d(k), n, _d(y), _d(m), LIK[likelihood], _d(out4), _d(lf))
    if info < 0:
        raise MemoryError("or_loglik_ext: allocation failed")
    return (out4[0], info, out4, lf) if factor else (out4[0], info, out4)


MEAN_KIND = {"zero": 0, "constant": 1, "linear": 2, "quadratic": 3}


def mean(kind, x, y):
    """(m, beta, xmin, ymin, info) of a mean function."""
    x = np.asfortranarray(x, dtype=np.float64)
    n, d = x.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = np.zeros(n)
    beta = np.zeros(d)
    xmin = np.zeros(d)
    ymin = ctypes.c_double(0.0)
    info = load().or_mean(MEAN_KIND[kind], n, d, _d(x), _d(y), _d(m), _d(beta), _d(xmin), ctypes.byref(ymin))
    return m, beta, xmin, ymin.value, info


class Baseline:
    """The reference's host algorithm (tensors + dgemv + LAPACK), OpenMP over particles."""

    def __init__(self, x, y, mean, spec: Spec, noise=1e-10, optimise_noise=False,
                 likelihood="marginal", lapack_path=None, prefix="scipy_"):
        x = np.asfortranarray(x, dtype=np.float64)
        n, d = x.shape
        self.spec = spec
        self._keep = (x, np.ascontiguousarray(y, dtype=np.float64),
                      None if mean is None else np.ascontiguousarray(mean, dtype=np.float64))
        self.h = load().or_baseline_create(n, d, _d(self._keep[0]), _d(self._keep[1]), _d(self._keep[2]),
                                           spec.ref(), noise, int(optimise_noise), LIK[likelihood],
                                           (lapack_path or "").encode(), prefix.encode())
        if not self.h:
            raise ValueError("baseline creation failed")
        self.p = nparams(spec, optimise_noise)

    @property
    def lapack(self) -> bool:
        return bool(load().or_baseline_lapack(self.h))

    def swarm(self, theta, threads=0):
        th = np.ascontiguousarray(np.atleast_2d(theta), dtype=np.float64)
        s = th.shape[0]
        ll = np.empty(s)
        info = np.empty(s, dtype=np.int32)
        load().or_baseline_swarm(self.h, s, _d(th), _d(ll), info.ctypes.data_as(_P_INT), threads)
        return ll, info

    def close(self):
        if self.h:
            load().or_baseline_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scipy_openblas_path():
    """Path of the OpenBLAS that scipy ships (LAPACK for the CPU baseline), or None."""
    try:
        import scipy
    except ImportError:
        return None
    libs = Path(scipy.__file__).resolve().parent.parent / "scipy.libs"
    cands = sorted(libs.glob("libscipy_openblas*.so")) if libs.exists() else []
    return str(cands[0]) if cands else None
