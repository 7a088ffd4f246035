"""Seeded synthetic disks for tests and benchmarks.

A disk is given as the four arrays the reference's InitializeFromFile route
reads (src_main/Init.c:86-105): surface density, radial and azimuthal
velocity, temperature, each NRAD*NSEC doubles in ring-major order.

Noise is drawn with SplitMix64: u(seed, k) = (splitmix64(seed ^ k) >> 11) * 2**-53
with k = i*NSEC + j for cells.  The radial grid follows FillPolar1DArrays
(SourceEuler.c:81-106) and is evaluated with Python's math module (the C
library), so the same seed gives the same bits on every host.

Profiles (code units, G = M_star = 1):
  Sigma  = sigma0 * Rmed**(-slope) * (1 + noise * (2u - 1))
  T      = h**2 / Rmed                      (so cs**2 = gamma*T, H/r = h)
  vrad   = 0
  vtheta = sqrt(1/Rmed) * sqrt(1 - T*Rmed)  (the pressure-corrected Keplerian
           velocity of InitGasVelocity, Pframeforce.c:716, with cs**2 = gamma*T)
"""
from __future__ import annotations

import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser of uint64 inputs (vectorised, wraps mod 2**64)."""
    z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, k: np.ndarray) -> np.ndarray:
    """u(seed, k) in [0, 1) with 53 random bits."""
    x = np.uint64(seed) ^ k.astype(np.uint64)
    return (splitmix64(x) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def radii(nrad: int, rmin: float, rmax: float, log_grid: bool = True) -> list[float]:
    """Radii[0..nrad] exactly as FillPolar1DArrays computes them."""
    if log_grid:
        lg = math.log(rmax / rmin)
        return [rmin * math.exp(float(i) / float(nrad) * lg) for i in range(nrad + 1)]
    d = (rmax - rmin) / float(nrad)
    return [rmin + d * float(i) for i in range(nrad + 1)]


def rmed(nrad: int, rmin: float, rmax: float, log_grid: bool = True) -> np.ndarray:
    r = radii(nrad, rmin, rmax, log_grid)
    out = np.empty(nrad)
    for i in range(nrad):
        a, b = r[i], r[i + 1]
        v = 2.0 / 3.0 * (b * b * b - a * a * a)
        out[i] = v / (b * b - a * a)
    return out


def seeded_disk(nrad: int, nsec: int, *, rmin: float = 0.4, rmax: float = 2.5, log_grid: bool = True,
                sigma0: float = 5e-4, slope: float = 0.5, h: float = 0.05, noise: float = 1e-3,
                seed: int = 12345) -> dict[str, np.ndarray]:
    """Returns dens, vrad, vtheta, temper as (nrad, nsec) float64 arrays."""
    rm = rmed(nrad, rmin, rmax, log_grid)
    sig_ring = np.array([sigma0 * math.pow(r, -slope) for r in rm])
    t_ring = np.array([h * h / r for r in rm])
    vt_ring = np.array([math.sqrt(1.0 / r) * math.sqrt(1.0 - (h * h / r) * r) for r in rm])
    k = np.arange(nrad * nsec, dtype=np.uint64).reshape(nrad, nsec)
    u = uniform(seed, k)
    dens = sig_ring[:, None] * (1.0 + noise * (2.0 * u - 1.0))
    temper = np.repeat(t_ring[:, None], nsec, axis=1)
    vtheta = np.repeat(vt_ring[:, None], nsec, axis=1)
    vrad = np.zeros((nrad, nsec))
    return {"dens": np.ascontiguousarray(dens), "vrad": vrad, "vtheta": np.ascontiguousarray(vtheta),
            "temper": np.ascontiguousarray(temper)}


def write_cfg(path: str, arr: np.ndarray) -> None:
    """ASCII file for InitializeFromFile: one value per line, ring-major,
    shortest repr that round-trips through the reference's %lf."""
    flat = np.ascontiguousarray(arr).ravel()
    with open(path, "w") as f:
        f.write("\n".join(repr(float(v)) for v in flat))
        f.write("\n")


def write_opacity_table(path: str, rows: int, cols: int, pgrid, tgrid, planck, rosseland) -> None:
    """The gas opacity table of OpacityTable SM03 in the layout its reader
    takes (InitOpacities, Opacity2.c:252-300): rows, cols, then the pressure
    grid, the temperature grid, the Planck and the Rosseland values, one
    number per line, each written so that %lf reads back the same double."""
    with open(path, "w") as f:
        f.write(f"{int(rows)}\n{int(cols)}\n")
        for a in (pgrid, tgrid, planck, rosseland):
            f.write("\n".join(repr(float(v)) for v in np.ravel(a)))
            f.write("\n")


# T2SI of fondam.h: code temperature unit in K (G Msun / (R/mu) / AU)
T2SI = 1.32712440018e20 / (8.3144598 / (2.4 * 0.001)) / 149597870700.0


def sor_disk(nrad: int = 2048, nsec: int = 3072, *, base_nrad: int = 128, base_nsec: int = 256,
             rmin: float = 0.4, rmax: float = 2.5, seed: int = 12345) -> dict[str, np.ndarray]:
    """The mixed-precision SOR workload's disk: the 128x256 seeded surface
    density interpolated bilinearly in (ln r, phi) (periodic in phi) onto the
    target log grid, T = 2000 K (r/0.4)^-2.513 (20 K at r = 2.5) in code
    units, vr = 0 and the pressure-corrected Keplerian vtheta of that T."""
    base = seeded_disk(base_nrad, base_nsec, rmin=rmin, rmax=rmax, seed=seed)
    lr_b = np.log(rmed(base_nrad, rmin, rmax, True))
    rm = rmed(nrad, rmin, rmax, True)
    lr = np.log(rm)
    # radial weights (clamped at the edges)
    fi = np.interp(lr, lr_b, np.arange(base_nrad, dtype=float))
    i0 = np.clip(np.floor(fi).astype(int), 0, base_nrad - 2)
    wi = (fi - i0)[:, None]
    # azimuthal weights on cell centres, periodic
    phi = (np.arange(nsec) + 0.5) / nsec
    fj = phi * base_nsec - 0.5
    j0 = np.floor(fj).astype(int)
    wj = (fj - j0)[None, :]
    j0m = j0 % base_nsec
    j1m = (j0 + 1) % base_nsec
    d = base["dens"]
    lo = d[i0][:, j0m] * (1 - wj) + d[i0][:, j1m] * wj
    hi = d[i0 + 1][:, j0m] * (1 - wj) + d[i0 + 1][:, j1m] * wj
    dens = lo * (1 - wi) + hi * wi
    t_ring = 2000.0 * (rm / 0.4) ** (-2.513) / T2SI
    vt_ring = np.sqrt(1.0 / rm) * np.sqrt(1.0 - t_ring * rm)
    return {"dens": np.ascontiguousarray(dens), "vrad": np.zeros((nrad, nsec)),
            "vtheta": np.ascontiguousarray(np.repeat(vt_ring[:, None], nsec, axis=1)),
            "temper": np.ascontiguousarray(np.repeat(t_ring[:, None], nsec, axis=1))}


# RHO2CGS of fondam.h (Msun*1000/(AU*100)^3): code density unit in g/cm^3
RHO2CGS = 1.98855e30 * 1000.0 / (149597870700.0 * 100.0) ** 3


def seeded_pebbles(disk: dict, *, ratio: float = 0.01, st_min: float = 0.01, st_max: float = 0.1,
                   bulkdens: float = 1.0, adiabind: float = 1.4, seed: int = 4242) -> dict[str, np.ndarray]:
    """The two-fluid workload's pebble disk on a gas disk of seeded_disk:
    surface density ratio * Sigma_gas, Stokes numbers log-uniform in
    [st_min, st_max] per ring (seeded), the grain sizes that give them with
    the Epstein law of PebbleStokesNumbers (tau = 2 rho_bulk s / Sigma /
    sqrt(gamma)), vr = 0 and the gas vtheta."""
    dens = disk["dens"]
    nrad, nsec = dens.shape
    u = uniform(seed, np.arange(nrad, dtype=np.uint64))
    st = np.exp(np.log(st_min) + u * (np.log(st_max) - np.log(st_min)))
    pbulk = bulkdens / RHO2CGS
    psize = st[:, None] * dens * math.sqrt(adiabind) / (2.0 * pbulk)
    return {"pdens": np.ascontiguousarray(ratio * dens), "pvrad": np.zeros((nrad, nsec)),
            "pvtheta": np.ascontiguousarray(disk["vtheta"].copy()), "psize": np.ascontiguousarray(psize),
            "stokes": st}
