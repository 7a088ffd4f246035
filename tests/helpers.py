"""Shared helpers of the test suite: golden fixtures, parameters, the planet
trajectory replay used to pin the hydro step independently of the N-body."""
from __future__ import annotations

import contextlib
import os
import tempfile

import numpy as np

from fargo_thorin_amd import Params, make_planets
from fargo_thorin_amd._capi import NBODY_FN, host_nbody_fn
from tests.golden.make_golden import CASES, FIELDS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name: str):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


_TABLE_DIR = None


def opacity_table_dir() -> str:
    """A directory holding the SM03 gas opacity table (the reference's
    data_gasopacity_malygin.txt, from the fixture sm03_gas_table.npz) in the
    layout its reader takes; made once per test session."""
    global _TABLE_DIR
    if _TABLE_DIR is None:
        import tempfile

        from fargo_thorin_amd.seed import write_opacity_table
        t = np.load(os.path.join(GOLDEN, "sm03_gas_table.npz"))
        _TABLE_DIR = tempfile.mkdtemp(prefix="thorin_sm03_")
        write_opacity_table(os.path.join(_TABLE_DIR, "data_gasopacity_malygin.txt"), t["rows"], t["cols"],
                            t["pgrid"], t["tgrid"], t["planck"], t["rosseland"])
    return _TABLE_DIR + "/"


def params_for(cfg: dict, **extra) -> Params:
    keys = {k: v for k, v in cfg.items() if not k.startswith("_")}
    keys.update({k: str(v) for k, v in extra.items()})
    keys.setdefault("OutputDir", "/nonexistent-thorin-out/")
    if str(keys.get("OpacityTable", ""))[:1] in "sS":
        keys.setdefault("InputDir", opacity_table_dir())
    return Params(**keys)


def planets_for(planets, cfg=None):
    """The planets of a case; with its config, the flags and the accretion
    rate every planet takes from it (ReboundInterface.c:191)."""
    feel = bool(cfg) and str(cfg.get("PlanetsFeelDisk", "NO"))[:1] in "yY"
    s = make_planets([m for m, a in planets], [a for m, a in planets], feel_disk=feel)
    acc = float(cfg.get("AccretionRate", 0.0)) if cfg else 0.0
    for k in range(s.nb):
        s.acc[k] = acc
    return s


class Replay:
    """N-body callback that moves the planets along the reference's own
    trajectory (tqwk0.dat rows: position at the start of each DT).  One
    substep per DT is assumed, which the golden runs have.  With integrate,
    the library's own N-body step advances the planets first, so that the
    velocities (which the log does not hold) follow the orbit to round-off;
    the positions are then replaced by the logged ones."""

    def __init__(self, tq, integrate: bool = False):
        self.tq = np.asarray(tq)
        self.n = 0
        self.cb = NBODY_FN(self._step)
        self.nb = host_nbody_fn() if integrate else None

    def _step(self, user, sysp, t, dt, angle):
        if self.nb is not None:
            rc = self.nb(user, sysp, t, dt, angle)
            if rc:
                return rc
        self.n += 1
        if self.n < len(self.tq):
            sysp.contents.x[0] = self.tq[self.n, 3]
            sysp.contents.y[0] = self.tq[self.n, 4]
        return 0


def case(name):
    cfg, planets, full = CASES[name]
    return cfg, planets, full


def rel_linf(a, b, floor=1e-300):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor))


def initial_disk(name):
    """The seeded initial disk of a golden case: the fixture's arrays where it
    keeps them, else made again from the seed with the case's disk keys."""
    from fargo_thorin_amd import seeded_disk
    g = golden(name)
    if "init_dens" in g:
        return {k: g["init_" + k] for k in ("dens", "vrad", "vtheta", "temper")}
    cfg, _, _ = case(name)
    return seeded_disk(int(cfg["Nrad"]), int(cfg["Nsec"]), log_grid=cfg.get("RadialSpacing", "LOG")[0] in "lL",
                       rmin=float(cfg["Rmin"]), rmax=float(cfg["Rmax"]), h=float(cfg["AspectRatio"]),
                       sigma0=float(cfg["Sigma0"]), slope=float(cfg["SigmaSlope"]))


def sha(a) -> np.ndarray:
    import hashlib
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@contextlib.contextmanager
def short_tmpdir():
    """A working directory for runs of the reference's driver, under /tmp:
    WriteDiskPolar forms its file names in an 80-byte buffer (Output.c:159),
    so the output directory's path has to stay short."""
    with tempfile.TemporaryDirectory(prefix="t", dir="/tmp") as d:
        yield d
