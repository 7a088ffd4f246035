"""The opacity laws on the GPU (thorin_opacity_eval: the device code the
energy kernels run) against the oracle's restatement of Opacity.c, which the
golden runs pin bit for bit against the reference.

Random (rho, T) over the disks' range, and for LP85 the points just either
side of its three density-dependent regime borders (T = k rho^e at relative
distances 1e-7 .. 3e-5): there the device decides the regime from a
single-precision estimate when T is clear of the border by more than 1e-5
and from the double expression otherwise, so every decision must be the
oracle's.  The laws are continuous at the borders, but a regime decided the
other way within 1e-5 of a border changes kappa by ~1e-6 relative, far
above the 1e-13 allowed here for the device's pow/exp/log (a few ulp).
"""
import ctypes as C

import numpy as np
import pytest

from fargo_thorin_amd import Disk, configs
from fargo_thorin_amd._capi import last_error, lib
from oracle import lib as olib
from tests.helpers import params_for

pytestmark = pytest.mark.gpu

LAWS = {"LP85": 1, "BL94": 0, "ZHU12": 2}
RNG = np.random.default_rng(20261204)


def _device(d, rho, T):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.empty_like(rho)
    dp = C.POINTER(C.c_double)
    rc = lib().thorin_opacity_eval(d._h, rho.ctypes.data_as(dp), T.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                   rho.size)
    assert rc == 0, last_error()
    return out


def _oracle(law, rho, T):
    f = olib().oracle_opacity_law
    return np.array([f(law, float(r), float(t), 0.0) for r, t in zip(rho, T)])


def _disk(table):
    cfg = configs.with_grid(configs.DAMPED_1024x1536, 16, 32, OpacityTable=table)
    return Disk(params_for(cfg))


@pytest.mark.parametrize("table", list(LAWS))
def test_opacity_law_on_device_matches_oracle(table):
    d = _disk(table)
    n = 20000
    rho = np.exp(RNG.uniform(np.log(1e-14), np.log(1e-5), n))
    T = np.exp(RNG.uniform(np.log(5.0), np.log(1e5), n))
    dev, ref = _device(d, rho, T), _oracle(LAWS[table], rho, T)
    rel = np.abs(dev - ref) / np.abs(ref)
    print(f"[gpu-opacity] {table}: max rel {rel.max():.2e} over {n} points", flush=True)
    assert rel.max() <= 1e-13, f"{table}: rho={rho[rel.argmax()]!r} T={T[rel.argmax()]!r}"


def test_lp85_regime_borders_decided_as_the_oracle():
    d = _disk("LP85")
    # T = k * rho^e: the borders between regimes 3|4, 5|6 and 6|7 (Opacity.c:214-220)
    borders = [(pow(5.0e-3 / 2.0e34, -1.0 / 10.0), 1.0 / 15.0),
               (pow(2.0e-8 / 1.0e-36, 1.0 / 7.0), 1.0 / 21.0),
               (pow(1.0e-36 / 1.5e20, -1.0 / 12.5), 4.0 / 75.0)]
    offs = np.array([1e-7, 3e-7, 1e-6, 3e-6, 1e-5, 3e-5])
    offs = np.concatenate([-offs[::-1], offs])
    rho, T = [], []
    for k, e in borders:
        for r in np.exp(RNG.uniform(np.log(1e-14), np.log(1e-5), 200)):
            tb = k * r ** e
            rho += [r] * offs.size
            T += list(tb * (1.0 + offs))
    rho, T = np.array(rho), np.array(T)
    dev, ref = _device(d, rho, T), _oracle(1, rho, T)
    rel = np.abs(dev - ref) / np.abs(ref)
    print(f"[gpu-opacity] LP85 borders: max rel {rel.max():.2e} over {rho.size} points", flush=True)
    assert rel.max() <= 1e-13, f"rho={rho[rel.argmax()]!r} T={T[rel.argmax()]!r}"
