"""The reference-style driver fargo_thorin_amd/thorin_amd (the library behind
main.c's loop: in.par, planet file, InitializeFromFile, SendOutput,
tqwk*.dat, used_rad.dat, restart with -s N) on the two-fluid golden case: its
gas*.dat and gaspebble*.dat files against the reference's own outputs, then a
restart from its output 1 (gas and pebble restart files, N-body state) against
the reference's output 2.  The driver integrates the planet with its own
N-body step (REBOUND in the reference), so fields are compared at 1e-10."""
import os
import subprocess

import numpy as np
import pytest

from fargo_thorin_amd import configs
from fargo_thorin_amd.seed import write_cfg
from tests.helpers import case, golden, rel_linf

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fargo_thorin_amd", "thorin_amd")
NAME = "pebbles_q3e-5_64x128"


def _read(outdir, name, n, shape):
    return np.fromfile(os.path.join(outdir, f"gas{name}{n}.dat"), dtype=np.float64).reshape(shape)


def test_gpu_driver_two_fluid_run_and_restart(tmp_path):
    print("[gpu-driver] two-fluid run", flush=True)
    cfg, planets, _ = case(NAME)
    g = golden(NAME)
    nr, ns = int(g["nrad"]), int(g["nsec"])
    cfgdir, outdir = tmp_path / "in", tmp_path / "out"
    cfgdir.mkdir()
    outdir.mkdir()
    for k in ("dens", "vrad", "vtheta", "temper"):
        write_cfg(str(cfgdir / f"gas{k}.cfg"), g[f"init_{k}"])
    pf = tmp_path / "planet.cfg"
    pf.write_text(configs.planet_text(planets))
    par = tmp_path / "in.par"
    par.write_text(configs.inpar_text(cfg, str(outdir) + "/", str(pf), str(cfgdir)))
    r = subprocess.run([DRIVER, str(par)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count(".") >= int(g["ntot"])
    for f in ("pebblesize", "pebblestokes", "pebbledens", "pebblevrad", "pebblevtheta"):
        assert np.array_equal(_read(outdir, f, 0, (nr, ns)), g[f"{f}0"]), f
    for n in (1, 2):
        for f in ("dens", "vrad", "vtheta", "temper", "pebbledens", "pebblevrad", "pebblevtheta"):
            key = f"{f}{n}"
            if key in g:
                err = rel_linf(_read(outdir, f, n, (nr, ns)), g[key])
                assert err <= 1e-10, f"{key}: {err:.3e}"
    used = np.loadtxt(outdir / "used_rad.dat")
    assert np.array_equal(used, g["used_rad"])
    tq = np.loadtxt(outdir / "tqwk0.dat", ndmin=2)
    assert np.array_equal(tq[:, 0], g["tqwk0"][:len(tq), 0])
    # the planet's position at the start of every DT: the library's N-body
    # step (with the disk's kicks and the pebble accretion) against REBOUND's
    err = np.max(np.abs(tq[:, 3:5] - g["tqwk0"][:len(tq), 3:5]))
    assert err <= 1e-10, f"planet trajectory: {err:.3e}"
    # restart from output 1 of this run; the reference's own restart (-s 1,
    # damping targets = ring means of the restart files) is the comparison
    print("[gpu-driver] restart -s 1", flush=True)
    for f in ("dens", "temper", "pebbledens", "pebblevrad", "pebblevtheta"):
        os.remove(outdir / f"gas{f}2.dat")
    r = subprocess.run([DRIVER, "-s", "1", str(par)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    for f in ("dens", "vrad", "vtheta", "temper", "pebbledens", "pebblevrad", "pebblevtheta"):
        err = rel_linf(_read(outdir, f, 2, (nr, ns)), g[f"restart_{f}2"])
        assert err <= 1e-10, f"restart {f}2: {err:.3e}"
    # nbody.orbits.dat of both runs (the set-up row, one row per DT, the
    # restart's row) against the reference's (ReboundInterface.c:405-441)
    orb = np.loadtxt(outdir / "nbody.orbits.dat", ndmin=2)
    ref = g["nbody_orbits"]
    assert orb.shape == ref.shape, f"{orb.shape} vs {ref.shape}"
    assert np.array_equal(orb[:, :2], ref[:, :2]), "planet index and time"
    for c in (2, 8, 9, 10, 11):           # a, mass, gas, solid, radius
        scale = max(np.max(np.abs(ref[:, c])), 1e-300)
        assert np.max(np.abs(orb[:, c] - ref[:, c])) <= 1e-10 * scale, f"column {c}"
    for c in (3, 4, 5) + tuple(range(12, 18)):   # e, inc, Omega, x .. vz
        assert np.max(np.abs(orb[:, c] - ref[:, c])) <= 1e-10, f"column {c}"
    lon = orb[:, 5] + orb[:, 6] + orb[:, 7] - ref[:, 5] - ref[:, 6] - ref[:, 7]
    assert np.max(np.abs((lon + np.pi) % (2 * np.pi) - np.pi)) <= 1e-9, "longitude"
