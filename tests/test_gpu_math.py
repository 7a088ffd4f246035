"""The device math library against the host's glibc, which the reference's
builds (and the oracle) call.  Each function is evaluated by
thorin_math_eval -- the device functions the kernels call, compiled with the
kernels' flags -- on the value ranges the hydro step feeds it, and by glibc
through ctypes; the distance in units in the last place (ulp) is bounded.

Where the kernels call them (fargo_thorin_amd/csrc/thorin_dev.hip):
  pow    opacity laws (T^b with |b| up to 24, rho^a), sound speed, scale
         heights, power-law profiles (r^e), MassTaper's sin^2
  exp    Flock alpha's and the opacity laws' exponentials
  log10  the SM03 table lookup (log P, log T)
  log    the grid's logarithmic spacing
  tanh   Flock alpha(T) (Viscosity.c:39-43)
  sin    the SM03 dust regime term and the irradiation angle
  sqrt   Keplerian speeds, sound speeds, distances
  atan2, asin, atan
         planet azimuths and the Hill-sphere sector bounds
A bound of 2 ulp holds when both libraries are faithfully rounded (each
within 1 ulp of the exact value); one call's 2 ulp is 4.4e-16 relative, six
orders below the 1e-10 field bar, so the differences these functions make
between the GPU and the reference stay within rounding of the step.
"""
import ctypes as C

import numpy as np
import pytest

from fargo_thorin_amd._capi import lib, last_error

pytestmark = pytest.mark.gpu

N = 1 << 16
RNG = np.random.default_rng(20261204)


def _logu(lo, hi, n):
    return np.exp(RNG.uniform(np.log(lo), np.log(hi), n))


def _glibc(name, x, y=None):
    libm = C.CDLL("libm.so.6")
    f = getattr(libm, name)
    f.restype = C.c_double
    f.argtypes = [C.c_double] * (1 if y is None else 2)
    if y is None:
        return np.array([f(float(a)) for a in x])
    return np.array([f(float(a), float(b)) for a, b in zip(x, y)])


def _device(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    dp = C.POINTER(C.c_double)
    yp = None if y is None else np.ascontiguousarray(y, dtype=np.float64).ctypes.data_as(dp)
    rc = lib().thorin_math_eval(fn, x.ctypes.data_as(dp), yp, out.ctypes.data_as(dp), x.size)
    assert rc == 0, last_error()
    return out


def _ulp(a, b):
    """Distance in ulp between finite doubles (0 for equal values)."""
    ia, ib = a.view(np.int64), b.view(np.int64)
    lo = np.int64(-0x8000000000000000)
    oa = np.where(ia < 0, lo - ia, ia)
    ob = np.where(ib < 0, lo - ib, ib)
    return np.abs(oa - ob)


# fn id, glibc name, x sampler, y sampler (two-argument functions)
CASES = {
    "pow_T": (0, "pow", lambda: _logu(3.0, 1e5, N), lambda: RNG.uniform(-24.0, 10.0, N)),
    "pow_rho": (0, "pow", lambda: _logu(1e-22, 1e-2, N), lambda: RNG.uniform(-1.0, 1.0, N)),
    "pow_r": (0, "pow", lambda: _logu(0.05, 60.0, N), lambda: RNG.uniform(-3.0, 3.0, N)),
    "exp": (1, "exp", lambda: RNG.uniform(-80.0, 80.0, N), None),
    "log": (2, "log", lambda: _logu(1e-30, 1e12, N), None),
    "log10": (3, "log10", lambda: _logu(1e-30, 1e12, N), None),
    "tanh": (4, "tanh", lambda: RNG.uniform(-25.0, 25.0, N), None),
    "sin": (5, "sin", lambda: RNG.uniform(-20.0, 20.0, N), None),
    "atan2": (6, "atan2", lambda: RNG.uniform(-40.0, 40.0, N), lambda: RNG.uniform(-40.0, 40.0, N)),
    "asin": (7, "asin", lambda: RNG.uniform(-1.0, 1.0, N), None),
    "atan": (8, "atan", lambda: RNG.uniform(-200.0, 200.0, N), None),
    "sqrt": (9, "sqrt", lambda: _logu(1e-30, 1e30, N), None),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_math_within_2_ulp_of_glibc(name):
    fn, cname, xs, ys = CASES[name]
    x = xs()
    y = ys() if ys is not None else None
    d = _device(fn, x, y)
    h = _glibc(cname, x, y)
    fin = np.isfinite(h)
    assert np.array_equal(np.isfinite(d), fin), f"{name}: finite/non-finite results differ"
    u = _ulp(d[fin], h[fin])
    hist = {int(k): int(v) for k, v in zip(*np.unique(np.minimum(u, 9), return_counts=True))}
    print(f"[gpu-math] {name}: ulp histogram {hist} over {fin.sum()} values", flush=True)
    assert u.max() <= 2, f"{name}: {u.max()} ulp at x={x[fin][u.argmax()]!r}"
    if name == "sqrt":
        # IEEE 754 square root is correctly rounded on both sides
        assert u.max() == 0
