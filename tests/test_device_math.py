"""CPU accuracy checks of the hand-written device math (runs without a GPU).

The covariance build evaluates every element with fb::exp_nonpos
(fb_common.h) from lengthscales given by fb::exp_cr, and the Cholesky
factorisation takes every pivot's square
root and its reciprocal from fb::goldschmidt_sqrt_rsqrt (fb_tiles.h) fed
by the hardware's reciprocal square root estimate.  Both are compiled for
the host from the same source (tests/native/fb_hostmath.hip) and compared
with long double truth (expl, sqrtl).  At the reference's 1e-10 nugget a
systematic bias of the build is what moves the log-likelihood (a 1 ulp bias
of every off-diagonal element acts like a 2e-16 change of the nugget), so
the mean signed error is checked as well as the largest.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

LIB = Path(__file__).resolve().parent / "native" / "libfb_hostmath.so"
LD = np.longdouble
P_DBL = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.fail(f"{LIB} missing: build it with `make -C tests/native` (build() does)")
    h = ctypes.CDLL(str(LIB))
    h.fbh_exp_nonpos.argtypes = [ctypes.c_int, P_DBL, P_DBL]
    h.fbh_exp_nonpos.restype = None
    h.fbh_exp_cr.argtypes = [ctypes.c_int, P_DBL, P_DBL]
    h.fbh_exp_cr.restype = None
    h.fbh_sqrt_rsqrt.argtypes = [ctypes.c_int, P_DBL, P_DBL, P_DBL, P_DBL]
    h.fbh_sqrt_rsqrt.restype = None
    return h


def _p(a):
    return a.ctypes.data_as(P_DBL)


def _exp(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib.fbh_exp_nonpos(len(x), _p(x), _p(y))
    return y


def _ulps(y, truth_ld):
    """Signed error of the FP64 values y against long double truth, in ulps
    of the FP64 result."""
    t64 = np.asarray(truth_ld, dtype=np.float64)
    return np.asarray((np.asarray(y, dtype=LD) - truth_ld) / np.asarray(np.spacing(np.abs(t64)), dtype=LD),
                      dtype=np.float64)


def test_exp_nonpos_accuracy_and_bias(lib):
    """Within 1 ulp of expl on [-708, 0] (normal results), mean signed error
    within 0.01 ulp: the range of covariance exponents, log-spaced from 1e-12
    to 708 and uniform on [-60, 0]."""
    rng = np.random.default_rng(0)
    x = np.concatenate([-np.exp(rng.uniform(np.log(1e-12), np.log(708.0), 200_000)),
                        -rng.uniform(0.0, 60.0, 200_000)])
    y = _exp(lib, x)
    err = _ulps(y, np.exp(x.astype(LD)))
    print(f"exp_nonpos: max |err| {np.abs(err).max():.3f} ulp, mean {err.mean():+.4f} ulp, "
          f"correctly rounded {np.mean(np.abs(err) <= 0.5):.4f}")
    assert np.abs(err).max() <= 1.0
    assert abs(err.mean()) <= 0.01
    assert np.mean(np.abs(err) <= 0.5) >= 0.95


def test_exp_nonpos_special_values(lib):
    """exp(0) = 1 exactly (the diagonal of every stationary term), the
    subnormal range within one subnormal step, underflow to +0 below it,
    -inf -> 0, NaN stays NaN, and |r| at the reduction boundaries."""
    y = _exp(lib, np.array([0.0, -0.0, -np.inf, np.nan, -1000.0, -746.0, -1e300]))
    assert y[0] == 1.0 and y[1] == 1.0
    assert y[2] == 0.0 and np.isnan(y[3]) and y[4] == 0.0 and y[5] == 0.0 and y[6] == 0.0
    xs = -np.linspace(708.5, 744.5, 2001)
    ys = _exp(lib, xs)
    truth = np.exp(xs.astype(LD))
    assert np.all(np.abs(np.asarray(ys, dtype=LD) - truth) <= LD(4.9406564584124654e-324))
    # arguments at odd multiples of ln2 / 2, where the reduced argument is largest
    k = np.arange(-1000, 0)
    xb = (k + 0.5) * np.log(2.0)
    xb = xb[xb >= -708.0]
    eb = _ulps(_exp(lib, xb), np.exp(xb.astype(LD)))
    assert np.abs(eb).max() <= 1.0


def test_goldschmidt_sqrt_rsqrt(lib):
    """Square root and its reciprocal over the pivot range (1e-300 .. 1e300)
    from estimates up to 2**-22 off (the hardware estimate is documented at
    2**29 ulp, 2**-23 relative): the square root correctly rounded (the
    residual correction), the reciprocal -- the multiplier of the column
    below the pivot -- within 2 ulp, neither biased."""
    rng = np.random.default_rng(1)
    n = 200_000
    x = np.exp(rng.uniform(np.log(1e-300), np.log(1e300), n))
    x[:1000] = np.exp(rng.uniform(np.log(1e-14), np.log(1e-6), 1000))    # pivots of a 1e-10 nugget
    r_true = 1.0 / np.sqrt(x.astype(LD))
    for rel in (0.0, 2.0 ** -22):
        y0 = np.asarray(r_true * (1 + LD(rel) * rng.uniform(-1, 1, n).astype(LD)), dtype=np.float64)
        sq, rsq = np.empty(n), np.empty(n)
        lib.fbh_sqrt_rsqrt(n, _p(x), _p(y0), _p(sq), _p(rsq))
        es = _ulps(sq, np.sqrt(x.astype(LD)))
        er = _ulps(rsq, r_true)
        print(f"estimate off by {rel:.1e}: sqrt max {np.abs(es).max():.3f} ulp (mean {es.mean():+.4f}, "
              f"correctly rounded {np.mean(np.abs(es) <= 0.5):.4f}); rsqrt max {np.abs(er).max():.3f} ulp "
              f"(mean {er.mean():+.4f})")
        assert np.abs(es).max() <= 0.5 and np.abs(er).max() <= 2.0
        assert abs(es.mean()) <= 0.01 and abs(er.mean()) <= 0.02


def test_exp_cr_matches_correct_rounding(lib):
    """The hyperparameter exp: within 0.5 ulp of expl everywhere on
    [-700, 700], equal to the correctly rounded value (expl rounded once to
    FP64) in >= 99.9 % of points, and equal to the C library's exp -- the
    reference's FP64 host exp -- in >= 99.9 % of the theta range."""
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-40, 40, 200_000), rng.uniform(-1, 1, 100_000),
                        rng.uniform(-700, 700, 50_000), [0.0, 1.0, -1.0, 0.5 * np.log(2.0)]])
    y = np.empty_like(x)
    lib.fbh_exp_cr(len(x), _p(x), _p(y))
    t = np.exp(x.astype(LD))
    err = _ulps(y, t)
    cr = np.asarray(t, dtype=np.float64)
    print(f"exp_cr: max |err| {np.abs(err).max():.4f} ulp, equal to the rounded long double value "
          f"{np.mean(y == cr):.5f}")
    assert np.abs(err).max() <= 0.5 + 1e-3
    assert np.mean(y == cr) >= 0.999
    assert y[-4] == 1.0
    libm = ctypes.CDLL("libm.so.6")
    libm.exp.restype = ctypes.c_double
    libm.exp.argtypes = [ctypes.c_double]
    g = np.array([libm.exp(v) for v in x[:50_000]])
    assert np.mean(g == y[:50_000]) >= 0.999
