"""ctypes binding of libferebus_hip.so (the C ABI in include/ferebus_hip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object has not been built.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parents[1]          # .../ferebus-v7_amd
LIB_PATH = PKG_ROOT / "lib" / "libferebus_hip.so"

FB_OK = 0
ERRORS = {-1: "FB_E_ARG", -2: "FB_E_SPEC", -3: "FB_E_HIP", -4: "FB_E_NOMEM",
          -5: "FB_E_NOGPU", -6: "FB_E_CAPACITY"}

# every entry point declared in include/ferebus_hip.h
EXPORTS = (
    "fb_create", "fb_destroy", "fb_nparams", "fb_set_stream", "fb_swarm_loglik",
    "fb_swarm_loglik_async", "fb_particle_loglik", "fb_get_alpha", "fb_swarm_loglik_grad",
    "fb_profile_enable", "fb_profile_read", "fb_factor_path", "fb_factor_groups", "fb_device_name",
    "fb_last_error",
    "fb_multi_create", "fb_multi_swarm_loglik", "fb_multi_destroy", "fb_multi_schedule",
    "fb_multi_pso_step",
    "fb_pso_create", "fb_pso_destroy", "fb_pso_init", "fb_pso_step", "fb_pso_retry", "fb_pso_select",
    "fb_pso_set_gbest", "fb_pso_positions", "fb_debug_matrix", "fb_swarm_loglik_static",
    "fb_pso_set_coefficients",
)


class FbError(RuntimeError):
    """A C ABI call returned a negative status."""

    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class KernelSpec(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int),
                ("node_type", ctypes.POINTER(ctypes.c_int)),
                ("node_ndim", ctypes.POINTER(ctypes.c_int)),
                ("active_dims", ctypes.POINTER(ctypes.c_int))]


_P_DBL = ctypes.POINTER(ctypes.c_double)
_P_INT = ctypes.POINTER(ctypes.c_int)
_lib = None


def library_path() -> Path:
    return Path(os.environ.get("FEREBUS_HIP_LIB", str(LIB_PATH)))


def load() -> ctypes.CDLL:
    """Loads the gfx950 library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(str(path))
    vp = ctypes.c_void_p
    sig = {
        "fb_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, _P_DBL, _P_DBL,
                                     _P_DBL, ctypes.POINTER(KernelSpec), ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "fb_destroy": (ctypes.c_int, [vp]),
        "fb_nparams": (ctypes.c_int, [vp, _P_INT]),
        "fb_set_stream": (ctypes.c_int, [vp, vp]),
        "fb_swarm_loglik": (ctypes.c_int, [vp, ctypes.c_int, _P_DBL, _P_DBL, _P_INT]),
        "fb_swarm_loglik_async": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp]),
        "fb_swarm_loglik_static": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp, vp]),
        "fb_particle_loglik": (ctypes.c_int, [vp, ctypes.c_int, _P_DBL, _P_DBL, _P_INT]),
        "fb_get_alpha": (ctypes.c_int, [vp, _P_DBL, _P_DBL, _P_DBL, _P_INT]),
        "fb_swarm_loglik_grad": (ctypes.c_int, [vp, ctypes.c_int, _P_DBL, _P_DBL, _P_DBL, _P_INT]),
        "fb_debug_matrix": (ctypes.c_int, [vp, _P_DBL, ctypes.c_int, _P_DBL, _P_INT]),
        "fb_profile_enable": (ctypes.c_int, [vp, ctypes.c_int]),
        "fb_factor_path": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "fb_factor_groups": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "fb_profile_read": (ctypes.c_int, [vp, ctypes.c_int, _P_DBL, ctypes.POINTER(ctypes.c_longlong)]),
        "fb_pso_create": (ctypes.c_int, [ctypes.POINTER(vp), vp, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double]),
        "fb_pso_destroy": (ctypes.c_int, [vp]),
        "fb_pso_init": (ctypes.c_int, [vp, _P_DBL, _P_DBL, _P_DBL, _P_INT]),
        "fb_pso_step": (ctypes.c_int, [vp, _P_DBL, _P_DBL, _P_DBL, _P_INT]),
        "fb_pso_retry": (ctypes.c_int, [vp, ctypes.c_int, _P_INT, _P_DBL, _P_DBL, _P_DBL, _P_INT]),
        "fb_pso_set_coefficients": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "fb_pso_select": (ctypes.c_int, [vp, _P_DBL, _P_DBL, _P_INT]),
        "fb_pso_set_gbest": (ctypes.c_int, [vp, _P_DBL, ctypes.c_double]),
        "fb_pso_positions": (ctypes.c_int, [vp, _P_DBL, _P_DBL]),
        "fb_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
        "fb_last_error": (ctypes.c_char_p, []),
        "fb_multi_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]),
        "fb_multi_swarm_loglik": (ctypes.c_int, [vp, _P_INT, ctypes.POINTER(_P_DBL), ctypes.POINTER(_P_DBL),
                                                 ctypes.POINTER(_P_INT)]),
        "fb_multi_destroy": (ctypes.c_int, [vp]),
        "fb_multi_schedule": (ctypes.c_int, [vp, _P_INT]),
        "fb_multi_pso_step": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(_P_DBL), ctypes.POINTER(_P_DBL),
                                             ctypes.POINTER(_P_DBL), ctypes.POINTER(_P_INT)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != FB_OK:
        msg = load().fb_last_error()
        raise FbError(code, where, msg.decode() if msg else "")


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_P_DBL)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_P_INT)


def device_name() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(256)
    check(lib.fb_device_name(buf, 256), "fb_device_name")
    return buf.value.decode()
