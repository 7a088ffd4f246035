"""The drop-in on the GPU: the reference's own driver (main.c, its parameter
reading, set-up, REBOUND, outputs and restarts) with AlgoGas replaced by
fargo_thorin_amd/shim/algogas_shim.c over libthorin_amd.so, run on the
inputs the golden fixtures were made from.  What it writes is compared with
what the reference wrote:

  every output field            1e-10 relative L-inf (whole fields or the
                                stored rings), SHA-256 of output 0
  SOR iterations (-d)           equal, substep by substep
  tqwk0.dat (UpdateLog)         time and position equal, accelerations and
                                torques to 1e-10 of the acceleration scale
  nbody.orbits.dat              every column to 1e-10 of its scale
  gas*999.dat, Crash! time      the first crash's dump to 1e-10, its time and
                                the crash marks equal
  restart (-s 1, -s 0)          the restarted run's output against the
                                reference's restarted run
"""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.golden.make_golden import CASES, FIELDS, prepare_case, run_args
from tests.helpers import golden, sha, short_tmpdir

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "thorin_gpu")
TOL = 1e-10

needs_bin = pytest.mark.skipif(not os.path.exists(BIN),
                               reason="oracle/_ref/thorin_gpu is built from the reference's sources (make -f oracle/ref.mk)")


@pytest.fixture
def run_dir():
    with short_tmpdir() as d:
        yield d


def _say(*a):
    print("[gpu-shim]", *a, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _run(tmp, args, timeout=600):
    r = subprocess.run([BIN] + args, cwd=tmp, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"thorin_gpu {args} exited {r.returncode}:\n{r.stdout[-3000:]}\n{r.stderr[-2000:]}"
    return r.stdout


def _field(outdir, fld, n, nrad, nsec):
    fn = os.path.join(outdir, f"gas{fld}{n}.dat")
    assert os.path.exists(fn), f"gas{fld}{n}.dat not written"
    a = np.fromfile(fn, dtype=np.float64)
    assert a.size == nrad * nsec, f"gas{fld}{n}.dat: {a.size} values"
    return a.reshape(nrad, nsec)


def _check_outputs(g, outdir):
    nrad, nsec = int(g["nrad"]), int(g["nsec"])
    stride = int(g["row_stride"]) if "row_stride" in g else None
    checked = 0
    for n in [int(x) for x in g["outputs"]]:
        for fld in FIELDS:
            if f"sha_{fld}{n}" not in g:
                continue
            a = _field(outdir, fld, n, nrad, nsec)
            if n == 0:
                # written by the reference's own set-up before any AlgoGas
                assert np.array_equal(sha(a), g[f"sha_{fld}0"]), f"{fld}0"
                continue
            ref = g[f"{fld}{n}"] if stride is None else g[f"rows_{fld}{n}"]
            got = a if stride is None else a[::stride]
            err = _rel(got, ref)
            assert err <= TOL, f"{fld}{n}: rel Linf {err:.3e}"
            rm = g[f"ringmean_{fld}{n}"]
            merr = float(np.max(np.abs(a.mean(axis=1) - rm)) / max(np.max(np.abs(rm)), np.max(np.abs(ref))))
            assert merr <= TOL, f"{fld}{n} ring means: {merr:.3e}"
            checked += 1
    assert checked > 0
    return checked


def _check_sor(g, stdout):
    niter = [int(m) for m in re.findall(r"SOR: Niter = (\d+)", stdout)]
    ref = list(g["sor_niter"])
    assert niter == ref, f"SOR iterations {niter[:8]}... vs {ref[:8]}..."


def _check_tqwk(g, outdir):
    tq = np.loadtxt(os.path.join(outdir, "tqwk0.dat"), ndmin=2)
    gt = g["tqwk0"]
    assert tq.shape == gt.shape, f"tqwk0.dat rows {tq.shape} vs {gt.shape}"
    # UpdateLog's row (Force.c:66): time, torque, normalised torque, x, y,
    # ax, ay.  Time and position exact; the accelerations (sums over the
    # grid in another order) and the torque to 1e-10 of |ax| + |ay| of the row
    for c in (0, 3, 4):
        assert np.array_equal(tq[:, c], gt[:, c]), f"tqwk0.dat column {c}"
    scale = np.abs(gt[:, 5]) + np.abs(gt[:, 6])
    for c in (1, 5, 6):
        bad = np.abs(tq[:, c] - gt[:, c]) > TOL * scale
        assert not bad.any(), f"tqwk0.dat column {c}, rows {np.nonzero(bad)[0][:5]}"
    nz = gt[:, 1] != 0.0
    lim = TOL * scale[nz] * np.abs(gt[nz, 2] / gt[nz, 1])
    assert np.all(np.abs(tq[nz, 2] - gt[nz, 2]) <= lim), "tqwk0.dat normalised torque"


def _check_orbits(g, outdir):
    if "nbody_orbits" not in g:
        return
    orb = np.loadtxt(os.path.join(outdir, "nbody.orbits.dat"), ndmin=2)
    ref = g["nbody_orbits"]
    assert orb.shape == ref.shape, f"nbody.orbits.dat {orb.shape} vs {ref.shape}"
    for c in range(1, ref.shape[1]):
        scale = max(np.max(np.abs(ref[:, c])), 1e-300)
        err = np.max(np.abs(orb[:, c] - ref[:, c])) / scale
        assert err <= TOL, f"nbody.orbits.dat column {c}: {err:.3e}"


@needs_bin
@pytest.mark.parametrize("name", ["wall_lp85_64x128", "damp_bl94_48x96", "flock_cavity_48x96", "std_nu_noind_48x96",
                                  "hill_q1e-4_48x96"])
def test_shim_runs_the_reference_driver(name, run_dir):
    cfg, planets, _ = CASES[name]
    g = golden(name)
    par, outdir, _, _ = prepare_case(name, cfg, planets, run_dir, binary=BIN)
    _say(name)
    out = _run(run_dir, run_args(cfg, par))
    n = _check_outputs(g, outdir)
    _check_sor(g, out)
    _check_tqwk(g, outdir)
    _check_orbits(g, outdir)
    _say(name, f"{n} output fields to {TOL}")


@needs_bin
def test_shim_two_fluid_planet_orbits_and_restart(run_dir):
    """Pebbles, drag back-reaction, a planet that feels the disk and accretes
    pebbles (the planet state goes through the library and REBOUND every
    substep); then the reference's restart from output 1 (-s 1), whose
    pebble damping targets come from the restart files."""
    name = "pebbles_q3e-5_64x128"
    cfg, planets, _ = CASES[name]
    g = golden(name)
    par, outdir, _, _ = prepare_case(name, cfg, planets, run_dir, binary=BIN)
    _say(name)
    out = _run(run_dir, run_args(cfg, par))
    _check_outputs(g, outdir)
    _check_sor(g, out)
    _say(name, "restart -s 1")
    _run(run_dir, ["-s", "1", par])
    # tqwk0.dat holds both runs' rows, as the reference's did
    _check_tqwk(g, outdir)
    nrad, nsec = int(g["nrad"]), int(g["nsec"])
    for fld in FIELDS:
        if f"restart_{fld}2" in g:
            err = _rel(_field(outdir, fld, 2, nrad, nsec), g[f"restart_{fld}2"])
            assert err <= TOL, f"restarted {fld}2: {err:.3e}"


@needs_bin
def test_shim_crash_dump(run_dir):
    name = "crash_cold_48x96"
    cfg, planets, _ = CASES[name]
    g = golden(name)
    par, outdir, _, _ = prepare_case(name, cfg, planets, run_dir, binary=BIN)
    _say(name)
    # as the fixture was made: a run under -d, then one without it for the
    # marks (tqwk0.dat and nbody.orbits.dat keep both runs' rows)
    _check_sor(g, _run(run_dir, run_args(cfg, par)))
    out = _run(run_dir, run_args(cfg, par, debug=False))
    _check_outputs(g, outdir)
    _check_tqwk(g, outdir)
    _check_orbits(g, outdir)
    nrad, nsec = int(g["nrad"]), int(g["nsec"])
    for fld in ("dens", "vrad", "vtheta", "temper"):
        err = _rel(_field(outdir, fld, 999, nrad, nsec), g[f"{fld}999"])
        assert err <= TOL, f"{fld}999: {err:.3e}"
    times = [float(x) for x in re.findall(r"Crash! at time (\S+)", out)]
    assert times == list(g["crash_time"])
    marks = "".join(m.group(1) for m in re.finditer(r"^([c.]+)(?:\t|$)", out, re.M))
    assert marks == bytes(g["crash_marks"]).decode()


@needs_bin
def test_shim_two_fluid_1024x1536_workload_through_restart(run_dir):
    """The two-fluid workload at its own size by the reference's route: a
    fresh run's output 0, the seeded pebble files, a restart from output 0
    (-s 0) and 100 DTs; rings every 64th to 1e-10, orbits, SOR iterations."""
    name = "pebbles_1024x1536_100"
    cfg, planets, _ = CASES[name]
    g = golden(name)
    _say(name, "set-up")
    par, outdir, _, fresh = prepare_case(name, cfg, planets, run_dir, binary=BIN)
    for k, v in fresh.items():
        assert np.array_equal(v, g[k]), f"fresh set-up {k}"
    _say(name, "100 DTs")
    out = _run(run_dir, run_args(cfg, par))
    _check_outputs(g, outdir)
    _check_sor(g, out)
    _check_tqwk(g, outdir)
    _check_orbits(g, outdir)
