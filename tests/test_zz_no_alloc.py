"""fb_swarm_loglik_static: the swarm evaluation with no allocation, every
work buffer in the library's per-device scratch array, against
fb_swarm_loglik (bit-identical: the same tile routines in the same order).

Two calls in a row on the same array, the second with a smaller input
(fewer particles and a smaller training set), then the first again: the
array keeps the last call's contents, which no call may depend on.  Device
buffers come from the HIP runtime the library itself uses (the process's
libamdhip64.so.7), through ctypes.
"""
import ctypes

import numpy as np
import pytest

from ferebus_amd import SwarmLikelihood, compile_kernel, fit_mean, synthetic
from ferebus_amd import _abi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")   # the runtime already loaded for the library
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def close(self):
        self.rt.hipDeviceSynchronize()
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _static(hip, gp, th):
    s = len(th)
    d_th = hip.put(np.ascontiguousarray(th, dtype=np.float64))
    d_ll, d_info = hip.alloc(8 * s), hip.alloc(4 * s)
    _abi.check(gp._lib.fb_swarm_loglik_static(gp._h, s, d_th, d_ll, d_info, None), "fb_swarm_loglik_static")
    return hip.get(d_ll, np.empty(s)), hip.get(d_info, np.empty(s, dtype=np.int32))


def _gp(n, S, seed, noise=1e-10):
    w = synthetic.water(n=n, seed=seed)
    prog = compile_kernel(w.kernel, w.kernels, 3)
    gp = SwarmLikelihood(w.x, w.y, prog, mean=fit_mean("constant", w.x, w.y), noise=noise, max_swarm=S)
    th = np.random.default_rng(seed).uniform(1e-3, 1.0, (S, gp.nparams))
    return gp, th


def test_static_matches_batched_two_calls(gpu):
    hip = _Hip()
    try:
        big, th_big = _gp(500, 12, seed=0)
        small, th_small = _gp(300, 5, seed=1, noise=1e-6)
        ref_big = big.swarm_loglik(th_big)
        ref_small = small.swarm_loglik(th_small)
        for gp, th, (rll, rinfo), label in ((big, th_big, ref_big, "N=500 S=12"),
                                            (small, th_small, ref_small, "N=300 S=5 (smaller, second)"),
                                            (big, th_big, ref_big, "N=500 S=12 again")):
            print(f"static call {label}", flush=True)
            ll, info = _static(hip, gp, th)
            assert np.array_equal(info, rinfo), label
            ok = info == 0
            assert ok.any()
            assert np.array_equal(ll[ok], rll[ok]), (label, np.abs(ll[ok] - rll[ok]).max())
        # a swarm beyond the array is refused before any work is enqueued
        w = synthetic.methanol(n=2000, seed=0)
        gp = SwarmLikelihood(w.x, w.y, compile_kernel(w.kernel, w.kernels, 12), mean="constant", max_swarm=4)
        d_th = hip.put(np.zeros((4, gp.nparams)))
        rc = gp._lib.fb_swarm_loglik_static(gp._h, 4, d_th, hip.alloc(32), hip.alloc(16), None)
        assert rc == -6, rc
        assert b"MiB" in gp._lib.fb_last_error()
        gp.close()
    finally:
        hip.close()
