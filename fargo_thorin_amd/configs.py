"""The benchmark and test workloads as in.par key sets.

Every key exists in the reference's registry (src_main/var.c:19-139) except
POWERLAWOPACITY, the build-only kappa = k0*T^2 option.

  plain_128x256    : 128x256 log grid on [0.4, 2.5], Sigma0 5e-4 r^-0.5, h 0.05,
                     alpha 1e-3, gamma 1.4, energy equation with the implicit
                     radiative diffusion (SOR), LP85 opacity, WALL inner
                     boundary, one 1e-13 planet at r = 1, DT 1e-3, 200 steps
  damped_1024x1536 : the same physics at 1024x1536 with damping zones at both
                     radial edges (ZEROVRAD), 100 steps -- the hydro-step
                     benchmark workload
  pebbles_1024x1536: the damped disk plus a pebble fluid (Sigma_p = 0.01
                     Sigma, Stokes numbers 0.01-0.1 per ring, seed.seeded_pebbles),
                     drag back-reaction, particle diffusion, a q = 3e-5 planet
                     at r = 1 that feels the disk and accretes pebbles, 500 steps
  ensemble_16      : 16 damped disks of uneven grids advanced together
                     (ensemble_configs): disk d draws u(1000+d, k), k = 0..3:
                     Nrad from {512, 640, 768, 896, 1024} (k=0), Nsec = 1.5 Nrad,
                     alpha log-uniform in [1e-4, 1e-2] (k=1), Sigma0 = 5e-4 x
                     10^(log10(0.3) + u) (k=2), h uniform in [0.03, 0.07] (k=3)
  sor_2048x3072    : the energy equation's implicit solve at 2048x3072 with
                     BL94 opacity on the interpolated disk of seed.sor_disk
                     (T 20-2000 K), dt = the hydro CFL step -- the solver
                     study workload
"""
from __future__ import annotations

BASE = {
    "Sigma0": "5e-4", "SigmaSlope": "0.5", "AspectRatio": "0.05", "FlaringIndex": "0.0",
    "AlphaViscosity": "1e-3", "Adiabind": "1.4", "EnergyEquation": "YES", "CoolingTime": "-1.0",
    "StellarIrradiation": "NO", "ThicknessSmoothing": "0.6", "Transport": "FARGO",
    "RadialSpacing": "LOG", "Rmin": "0.4", "Rmax": "2.5", "OpacityTable": "LP85",
    "PlanetsFeelDisk": "NO", "WriteTorqueFiles": "YES", "DT": "1e-3", "InitializeFromFile": "YES",
    "Frame": "FIXED", "OmegaFrame": "0.0", "IndirectTerm": "YES",
}

PLAIN_128x256 = dict(BASE, Nrad="128", Nsec="256", InnerBoundary="WALL", Ntot="200", Ninterm="200")

DAMPED_1024x1536 = dict(BASE, Nrad="1024", Nsec="1536", InnerBoundary="DAMPING", DampTowards="ZEROVRAD",
                        Ntot="100", Ninterm="100")

PEBBLES_1024x1536 = dict(DAMPED_1024x1536, PebbleAccretion="YES", BackReaction="YES", PlanetsFeelDisk="YES",
                         Ntot="500", Ninterm="500")
PLANET_Q3E5 = [(3e-5, 1.0)]

SOR_2048x3072 = dict(BASE, Nrad="2048", Nsec="3072", InnerBoundary="DAMPING", DampTowards="ZEROVRAD",
                     OpacityTable="BL94", Ntot="1", Ninterm="1")

PLANET_ZERO_MASS = [(1e-13, 1.0)]


def with_grid(cfg: dict, nrad: int, nsec: int, **extra) -> dict:
    out = dict(cfg, Nrad=str(nrad), Nsec=str(nsec))
    out.update({k: str(v) for k, v in extra.items()})
    return out


def inpar_text(cfg: dict, outputdir: str, planetfile: str, cfgdir: str) -> str:
    """An in.par for the reference's reader (Interpret.c:84-119)."""
    lines = [f"{k}\t{v}" for k, v in cfg.items()]
    lines += [f"OutputDir\t{outputdir}", f"PlanetConfig\t{planetfile}",
              f"DensInfile\t{cfgdir}/gasdens.cfg", f"VradInfile\t{cfgdir}/gasvrad.cfg",
              f"VthetaInfile\t{cfgdir}/gasvtheta.cfg", f"TemperInfile\t{cfgdir}/gastemper.cfg"]
    return "\n".join(lines) + "\n"


def planet_text(planets) -> str:
    """PlanetConfig lines: name, mass, gas, solid, a, e, inc, Omega, omega, f
    (ReboundInterface.c:157-166)."""
    return "".join(f"Planet{k + 1}\t{m!r}\t0.0\t0.0\t{a!r}\t0.0\t0.0\t0.0\t0.0\t0.0\n"
                   for k, (m, a) in enumerate(planets))


ENSEMBLE_SIZES = [512, 640, 768, 896, 1024]


def ensemble_configs(n: int = 16, seed0: int = 1000, sizes=None) -> list[dict]:
    """The ensemble's members: in.par keys plus the seed of each disk (key
    "_seed"; it is not an in.par key).  sizes replaces the Nrad choices (for
    small test ensembles); Nsec = 1.5 Nrad."""
    import math

    import numpy as np

    from .seed import uniform
    sizes = ENSEMBLE_SIZES if sizes is None else list(sizes)
    out = []
    for d in range(n):
        seed = seed0 + d
        u = uniform(seed, np.arange(4, dtype=np.uint64))
        nrad = sizes[min(int(u[0] * len(sizes)), len(sizes) - 1)]
        nsec = nrad * 3 // 2
        alpha = 10.0 ** (-4.0 + 2.0 * u[1])
        sigma0 = 5e-4 * 10.0 ** (math.log10(0.3) + u[2])
        h = 0.03 + 0.04 * u[3]
        cfg = with_grid(DAMPED_1024x1536, nrad, nsec, AlphaViscosity=repr(float(alpha)),
                        Sigma0=repr(float(sigma0)), AspectRatio=repr(float(h)))
        cfg["_seed"] = seed
        out.append(cfg)
    return out
