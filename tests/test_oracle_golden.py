"""The CPU oracle against the reference's own outputs (tests/golden/).

The fixtures were written by tests/golden/make_golden.py, which runs the
reference FARGO_THORIN (built from its sources by oracle/ref.mk) on seeded
disks.  The oracle must reproduce every output field bit for bit, the SOR
iteration count of every substep and PhysicalTime of every DT.  The planets
follow the reference's own trajectory (Replay), so that the comparison pins
the hydro step and not the N-body integrator.
"""
import hashlib

import numpy as np
import pytest

from oracle import OracleDisk
from tests.helpers import FIELDS, Replay, case, golden, params_for, planets_for

FULL_CASES = ["wall_lp85_64x128", "damp_bl94_48x96", "dampinit_zhu12_40x64", "planet_q3e-5_48x96", "irr_lp85_48x96",
              "damp_sm03_48x96", "gasacc_q1e-3_64x128", "flock_cavity_48x96", "dampflux_48x96", "drift_48x96",
              "cold_lp85_48x96", "std_nu_noind_48x96", "hill_q1e-4_48x96"]


def run_oracle_case(name):
    cfg, planets, _ = case(name)
    g = golden(name)
    o = OracleDisk(params_for(cfg))
    o.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sysp = planets_for(planets, cfg)
    rep = Replay(g["tqwk0"])
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    snaps = {0: {f: o.field(f) for f in FIELDS}}
    niters, times = [], [o.state().physical_time]
    for i in range(ntot):
        lg = o.algogas(sysp, rep.cb)
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        times.append(o.state().physical_time)
        if (i + 1) % ninterm == 0:
            snaps[(i + 1) // ninterm] = {f: o.field(f) for f in FIELDS}
    return g, snaps, niters, times


@pytest.mark.parametrize("name", FULL_CASES)
def test_oracle_bitwise_fields(name):
    g, snaps, niters, times = run_oracle_case(name)
    checked = 0
    for n, fields in snaps.items():
        for f, a in fields.items():
            key = f"{f}{n}"
            if key not in g:
                continue
            b = g[key]
            assert np.array_equal(a, b), f"{key}: {np.count_nonzero(a != b)} cells differ"
            checked += 1
    assert checked >= 8


@pytest.mark.parametrize("name", FULL_CASES)
def test_oracle_sor_iterations_and_time(name):
    g, snaps, niters, times = run_oracle_case(name)
    ntot = int(g["ntot"])
    ref_niter = list(g["sor_niter"][:len(niters)])
    assert niters == ref_niter
    # the reference logs PhysicalTime at the start of every DT (tqwk0.dat)
    ref_t = g["tqwk0"][:ntot, 0]
    assert np.array_equal(np.array(times[:ntot]), ref_t)


def test_oracle_plain_128x256_200_steps_checksums():
    """The 128x256 workload for its full 200 steps: every output field's
    SHA-256 equals the reference's."""
    name = "plain_128x256_200"
    cfg, planets, _ = case(name)
    g = golden(name)
    o = OracleDisk(params_for(cfg))
    o.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sysp = planets_for(planets)
    rep = Replay(g["tqwk0"])
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    for i in range(ntot + 1):
        n = i // ninterm
        if i % ninterm == 0:
            for f in FIELDS:
                key = f"sha_{f}{n}"
                if key in g:
                    h = np.frombuffer(hashlib.sha256(o.field(f).tobytes()).digest(), dtype=np.uint8)
                    assert np.array_equal(h, g[key]), f"{f} at output {n}"
        if i == ntot:
            break
        o.algogas(sysp, rep.cb)


PEB_CASE = "pebbles_q3e-5_64x128"
PEB_FIELDS = ["pebbledens", "pebblevrad", "pebblevtheta"]


@pytest.mark.parametrize("name", [PEB_CASE, "pebheat_q3e-5_64x128"])
def test_oracle_pebbles_bitwise(name):
    """Two-fluid run with a q = 3e-5 planet that feels the disk and accretes
    pebbles: the equilibrium pebble disk of a fresh run (sizes, Stokes
    numbers, density, velocities), then every gas and pebble output field,
    the SOR iterations and PhysicalTime equal the reference's bit for bit."""
    cfg, planets, _ = case(name)
    g = golden(name)
    o = OracleDisk(params_for(cfg))
    o.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    o.init_pebbles(0)
    for f in PEB_FIELDS + ["pebblesize", "pebblestokes"]:
        assert np.array_equal(o.field(f), g[f"{f}0"]), f"initial {f}"
    sysp = planets_for(planets, cfg)
    m0 = sysp.mass[0]
    rep = Replay(g["tqwk0"], integrate=True)
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    niters, times = [], [o.state().physical_time]
    checked = 0
    for i in range(ntot):
        lg = o.algogas(sysp, rep.cb)
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        times.append(o.state().physical_time)
        if (i + 1) % ninterm == 0:
            n = (i + 1) // ninterm
            for f in FIELDS + PEB_FIELDS:
                key = f"{f}{n}"
                if key in g:
                    a = o.field(f)
                    assert np.array_equal(a, g[key]), f"{key}: {np.count_nonzero(a != g[key])} cells differ"
                    checked += 1
    assert checked >= 8
    assert niters == list(g["sor_niter"][:len(niters)])
    assert np.array_equal(np.array(times[:ntot]), g["tqwk0"][:ntot, 0])
    # the planet accreted pebbles
    assert sysp.mass[0] > m0 and sysp.solid[0] > 0.0


def test_oracle_crash_run_and_dump():
    """A cold disk with one cell at -0.01 of its temperature: DetectCrash
    fires in the first two DTs (the reference's marks 'cc..'), the first
    crash's dump (gas{dens,vrad,vtheta,temper}999.dat) and its time, then
    every field after three DTs, all bit for bit."""
    name = "crash_cold_48x96"
    cfg, planets, _ = case(name)
    g = golden(name)
    o = OracleDisk(params_for(cfg))
    o.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sysp = planets_for(planets, cfg)
    rep = Replay(g["tqwk0"])
    marks, dumps = "", []
    for i in range(int(g["ntot"])):
        lg = o.algogas(sysp, rep.cb)
        marks += "".join("c" if lg.crashed[k] else "." for k in range(lg.nsub))
        if lg.crash_dumped:
            dumps.append(lg.crash_time)
    assert marks == bytes(g["crash_marks"]).decode()[: len(marks)]
    assert dumps == list(g["crash_time"])
    for f in ("dens", "vrad", "vtheta", "temper"):
        assert np.array_equal(o.field("crash_" + f), g[f"{f}999"]), f"{f}999"
    for f in ("dens", "vrad", "vtheta", "temper", "energy"):
        assert np.array_equal(o.field(f), g[f"{f}1"]), f"{f}1"
