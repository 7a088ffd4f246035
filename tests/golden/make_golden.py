#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ by running the reference.

The reference FARGO_THORIN is compiled from its own sources by oracle/ref.mk
into oracle/_ref/thorin (sequential build).  This script runs it on seeded
disks (fargo_thorin_amd.seed) through its InitializeFromFile route and keeps
what it writes: the gas*<n>.dat fields, tqwk0.dat (PhysicalTime per DT),
used_rad.dat and the per-DT substep marks printed on stdout.  Only data is
stored; no reference source is copied.

    python tests/golden/make_golden.py            (needs oracle/_ref/thorin)
"""
from __future__ import annotations

import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fargo_thorin_amd import configs  # noqa: E402
from fargo_thorin_amd.seed import seeded_disk, write_cfg, write_opacity_table  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "thorin")
OUT = os.path.join(ROOT, "tests", "golden")
# the reference's gas opacity table of OpacityTable SM03 (data), kept as the
# fixture sm03_gas_table.npz
MALYGIN = "/root/reference/in/data_gasopacity_malygin.txt"
TABLE = os.path.join(OUT, "sm03_gas_table.npz")


def gas_table():
    if not os.path.exists(TABLE):
        tok = open(MALYGIN).read().split()
        rows, cols = int(tok[0]), int(tok[1])
        v = np.array([float(x) for x in tok[2:2 + rows + cols + 2 * rows * cols]])
        np.savez_compressed(TABLE, rows=rows, cols=cols, pgrid=v[:rows], tgrid=v[rows:rows + cols],
                            planck=v[rows + cols:rows + cols + rows * cols],
                            rosseland=v[rows + cols + rows * cols:])
    return np.load(TABLE)

FIELDS = ["dens", "vrad", "vtheta", "temper", "energy", "divv", "qplus", "opacity",
          "pebbledens", "pebblevrad", "pebblevtheta", "pebblesize", "pebblestokes"]

# name: (config keys, planets, stored fields).  Stored fields: True keeps every
# output field whole; False keeps every (nrad//8)-th ring; an int k keeps every
# k-th ring.  Ring means and SHA-256 digests are kept for every field.  A config
# key "_seeded_pebbles" (not an in.par key) runs the two-fluid workload's seeded
# pebble disk through the reference's restart route: a fresh run writes output
# 0 (gas fields and the N-body archive), the seeded pebble fields replace its
# gaspebble*0.dat files, and the run restarts from output 0 with -s 0.
CASES = {
    # config-1 physics on a small grid, WALL boundary, LP85 opacity
    "wall_lp85_64x128": (configs.with_grid(configs.PLAIN_128x256, 64, 128, Ntot=12, Ninterm=6,
                                           WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                           WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # damping zones at both edges (ZEROVRAD), Bell & Lin opacity
    "damp_bl94_48x96": (configs.with_grid(configs.DAMPED_1024x1536, 48, 96, Ntot=10, Ninterm=5, OpacityTable="BL94",
                                          WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                          WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # damping towards the initial state, Zhu et al. 2012 opacity, arithmetic grid
    "dampinit_zhu12_40x64": (configs.with_grid(configs.DAMPED_1024x1536, 40, 64, Ntot=8, Ninterm=4,
                                               OpacityTable="ZHU12", DampTowards="INIT", RadialSpacing="ARITHMETIC",
                                               WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                               WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # a 10 Earth-mass planet (q = 3e-5) that does not feel the disk
    "planet_q3e-5_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=8, Ninterm=4, WriteEnergy="YES"),
                           [(3e-5, 1.0)], True),
    # two-fluid disk: equilibrium pebble disk of a fresh run (coagulation-drift
    # model), drag back-reaction, particle diffusion, a q = 3e-5 planet that
    # feels the disk and accretes pebbles, evaporation above 1000 K
    "pebbles_q3e-5_64x128": (configs.with_grid(configs.DAMPED_1024x1536, 64, 128, Ntot=10, Ninterm=5,
                                               PebbleAccretion="YES", BackReaction="YES",
                                               PlanetsFeelDisk="YES", WriteEnergy="YES"),
                             [(3e-5, 1.0)], True),
    # stellar irradiation (flaring surface, Qirr in the energy equation's RHS)
    "irr_lp85_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=8, Ninterm=4, StellarIrradiation="YES",
                                         WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES", WriteOpacity="YES"),
                       configs.PLANET_ZERO_MASS, True),
    # Semenov et al. 2003 dust + Malygin et al. 2014 gas opacity (OpacityTable SM03),
    # damping zones, T 20-2000 K so that the dust regimes and the gas table both act
    "damp_sm03_48x96": (configs.with_grid(configs.DAMPED_1024x1536, 48, 96, Ntot=8, Ninterm=4, OpacityTable="SM03",
                                          WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                          WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # the two-fluid run with accretion heating from the pebble and the gas
    # accretion (AccretionalHeating, GasAccretionHeating; AccretionRate 5)
    "pebheat_q3e-5_64x128": (configs.with_grid(configs.DAMPED_1024x1536, 64, 128, Ntot=10, Ninterm=5,
                                               PebbleAccretion="YES", BackReaction="YES", PlanetsFeelDisk="YES",
                                               AccretionalHeating="YES", GasAccretionHeating="YES",
                                               AccretionRate="5.0", HeatingDelay="1", WriteEnergy="YES"),
                             [(3e-5, 1.0)], True),
    # gas accretion onto a Jupiter-mass planet (Kley's scheme, AccretionRate 5)
    "gasacc_q1e-3_64x128": (configs.with_grid(configs.PLAIN_128x256, 64, 128, Ntot=8, Ninterm=4, AccretionRate="5.0",
                                              WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                              WriteOpacity="YES"), [(1e-3, 1.0)], True),
    # the 128x256 / 200-step workload itself: checksums and ring profiles only
    "plain_128x256_200": (dict(configs.PLAIN_128x256, Ninterm="100", WriteEnergy="YES"),
                          configs.PLANET_ZERO_MASS, False),
    # the 1024x1536 damped workload for its 100 steps (outputs after 50 and 100)
    "damped_1024x1536_100": (dict(configs.DAMPED_1024x1536, Ninterm="50", WriteEnergy="YES"),
                             configs.PLANET_ZERO_MASS, 64),
    # the 1024x1536 two-fluid workload (seeded pebble disk, Sigma_p = 0.01 Sigma,
    # Stokes numbers 0.01-0.1 per ring), its first 100 of 500 steps
    "pebbles_1024x1536_100": (dict(configs.PEBBLES_1024x1536, Ntot="100", Ninterm="50", WriteEnergy="YES",
                                   _seeded_pebbles="1"), configs.PLANET_Q3E5, 64),
    # Flock et al. (2016) alpha(T) (the MRI transition at 900 K lies inside the
    # disk) and the viscosity cavity with its time-dependent LambdaDoubling scale
    "flock_cavity_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=8, Ninterm=4, AlphaFlock="YES",
                                             TMRI="900.0", TWidth="100.0", AlphaMRI="1e-2", AlphaDead="1e-4",
                                             CavityRadius="0.8", CavityRatio="0.1", CavityWidth="2.0",
                                             LambdaDoubling="50.0", WriteEnergy="YES", WriteDivV="YES",
                                             WriteQplus="YES", WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # damping zones towards a prescribed radial mass flux (DampTowards G, GasFlux)
    "dampflux_48x96": (configs.with_grid(configs.DAMPED_1024x1536, 48, 96, Ntot=8, Ninterm=4, DampTowards="G",
                                         GasFlux="1e-6", WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                         WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # LP85 on a disk colder than the ice line (166.8 K) everywhere: only the
    # ice-grain branch kappa = 2e-4 T^2 acts, which POWERLAW with
    # PowerLawOpacity 2e-4 reproduces
    "cold_lp85_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=8, Ninterm=4, Rmin="1.2",
                                          AspectRatio="0.025", WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                          WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # the cold disk with one cell at -0.01 of its temperature (negative
    # energy, positive density): DetectCrash fires in DTs 0 and 1, the fields
    # stay finite and the disk recovers in DT 2
    "crash_cold_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=3, Ninterm=3, Rmin="1.2",
                                           AspectRatio="0.025", WriteEnergy="YES", _temper_cell="20,40,-0.01"),
                         configs.PLANET_ZERO_MASS, True),
    # the standard (non-FARGO) azimuthal transport, a constant kinematic
    # viscosity instead of alpha, and no indirect term
    "std_nu_noind_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=8, Ninterm=4, Transport="STANDARD",
                                             AlphaViscosity="0.0", Viscosity="1e-5", IndirectTerm="NO",
                                             WriteEnergy="YES", WriteDivV="YES", WriteQplus="YES",
                                             WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
    # a q = 1e-4 planet that feels the disk with its Hill sphere excluded
    # from the disk force (ExcludeHill, HillCut 0.8)
    "hill_q1e-4_48x96": (configs.with_grid(configs.PLAIN_128x256, 48, 96, Ntot=8, Ninterm=4, ExcludeHill="YES",
                                           PlanetsFeelDisk="YES", WriteEnergy="YES", WriteDivV="YES",
                                           WriteQplus="YES", WriteOpacity="YES"), [(1e-4, 1.0)], True),
    # an imposed supplementary torque that drives the disk inwards (ImposedDiskDrift)
    "drift_48x96": (configs.with_grid(configs.DAMPED_1024x1536, 48, 96, Ntot=8, Ninterm=4,
                                      ImposedDiskDrift="-1e-4", WriteEnergy="YES", WriteDivV="YES",
                                      WriteQplus="YES", WriteOpacity="YES"), configs.PLANET_ZERO_MASS, True),
}


def _write_raw(path, arr):
    np.ascontiguousarray(arr, dtype=np.float64).tofile(path)


def prepare_case(name, cfg, planets, tmp, binary=REF):
    """Writes a case's inputs under tmp (seeded InitializeFromFile arrays, the
    opacity table, planet.cfg, in.par) and returns (par, outdir, disk, extra).
    A "_seeded_pebbles" case first runs `binary` fresh for output 0 and puts
    the seeded pebble fields in its place (extra: the fresh set-up's SHA-256
    digests); the run then restarts from output 0 (-s 0)."""
    cfg = dict(cfg)
    seeded_pebbles = bool(cfg.pop("_seeded_pebbles", ""))
    # "_temper_cell": "i,j,factor" scales one cell of the initial temperature
    temper_cell = cfg.pop("_temper_cell", "")
    nrad, nsec = int(cfg["Nrad"]), int(cfg["Nsec"])
    cfgdir = os.path.join(tmp, "in")
    outdir = os.path.join(tmp, "out")
    os.makedirs(cfgdir)
    os.makedirs(outdir)
    disk = seeded_disk(nrad, nsec, log_grid=cfg.get("RadialSpacing", "LOG")[0] in "lL",
                       rmin=float(cfg["Rmin"]), rmax=float(cfg["Rmax"]), h=float(cfg["AspectRatio"]),
                       sigma0=float(cfg["Sigma0"]), slope=float(cfg["SigmaSlope"]))
    if temper_cell:
        ci, cj, cf = temper_cell.split(",")
        disk["temper"] = disk["temper"].copy()
        disk["temper"][int(ci), int(cj)] *= float(cf)
    for k in ("dens", "vrad", "vtheta", "temper"):
        write_cfg(os.path.join(cfgdir, f"gas{k}.cfg"), disk[k])
    if cfg.get("OpacityTable", "")[:1] in "sS":
        t = gas_table()
        write_opacity_table(os.path.join(cfgdir, "data_gasopacity_malygin.txt"), t["rows"], t["cols"],
                            t["pgrid"], t["tgrid"], t["planck"], t["rosseland"])
    pf = os.path.join(tmp, "planet.cfg")
    with open(pf, "w") as f:
        f.write(configs.planet_text(planets))
    par = os.path.join(tmp, "in.par")
    extra = {}
    if seeded_pebbles:
        # fresh run that writes output 0 only (and one DT that is discarded)
        with open(par, "w") as f:
            f.write(configs.inpar_text(dict(cfg, Ntot="0"), outdir + "/", pf, cfgdir))
        r0 = subprocess.run([binary, par], cwd=tmp, capture_output=True, text=True, timeout=3600)
        if r0.returncode != 0:
            raise RuntimeError(f"set-up run failed on {name}:\n{r0.stdout[-2000:]}\n{r0.stderr[-2000:]}")
        for k in ("dens", "vrad", "vtheta", "temper"):
            a = np.fromfile(os.path.join(outdir, f"gas{k}0.dat"), dtype=np.float64)
            extra[f"fresh_sha_{k}0"] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), dtype=np.uint8)
        pb = seeded_pebbles_of(disk, cfg)
        for k, src in (("dens", "pdens"), ("vrad", "pvrad"), ("vtheta", "pvtheta"), ("size", "psize")):
            _write_raw(os.path.join(outdir, f"gaspebble{k}0.dat"), pb[src])
        for fn in os.listdir(outdir):
            if fn.startswith("tqwk"):
                os.remove(os.path.join(outdir, fn))
    with open(par, "w") as f:
        f.write(configs.inpar_text(cfg, outdir + "/", pf, cfgdir))
    return par, outdir, disk, extra


def run_args(cfg, par, debug=True):
    """The command-line arguments of a case's main run."""
    return (["-d"] if debug else []) + (["-s", "0"] if cfg.get("_seeded_pebbles") else []) + [par]


def run_case(name, cfg, planets, full):
    nrad, nsec = int(cfg["Nrad"]), int(cfg["Nsec"])
    with tempfile.TemporaryDirectory() as tmp:
        par, outdir, disk, restart0 = prepare_case(name, cfg, planets, tmp)
        # -d: the reference prints its CFL block and "SOR: Niter = n" for every substep
        cmd = [REF] + run_args(cfg, par)
        r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=7200)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed on {name}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        ntot, ninterm = int(cfg["Ntot"]), int(cfg["Ninterm"])
        outs = list(range(0, ntot // ninterm + 1))
        data = {"nrad": nrad, "nsec": nsec, "ntot": ntot, "ninterm": ninterm, "outputs": np.array(outs)}
        data.update(restart0)
        if not isinstance(full, int) or isinstance(full, bool):
            # small grids keep the seeded initial arrays; the large ones are
            # made again from the seed (fargo_thorin_amd.seed.seeded_disk)
            for k in ("dens", "vrad", "vtheta", "temper"):
                data["init_" + k] = disk[k]
        stride = None if full is True else (max(1, nrad // 8) if full is False else int(full))
        for n in outs:
            for fld in FIELDS:
                fn = os.path.join(outdir, f"gas{fld}{n}.dat")
                if not os.path.exists(fn):
                    continue
                a = np.fromfile(fn, dtype=np.float64).reshape(nrad, nsec)
                data[f"sha_{fld}{n}"] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), dtype=np.uint8)
                data[f"ringmean_{fld}{n}"] = a.mean(axis=1)
                if stride is None:
                    data[f"{fld}{n}"] = a
                elif n > 0 or full is False:
                    data[f"rows_{fld}{n}"] = a[::stride].copy()
        if name == "pebbles_q3e-5_64x128":
            # restart from output 1 (-s 1): the pebble damping targets become
            # the ring means of the restart files (BckpFieldsForBC); keep the
            # output 2 the restarted run writes
            r2 = subprocess.run([REF, "-s", "1", par], cwd=tmp, capture_output=True, text=True, timeout=3600)
            if r2.returncode != 0:
                raise RuntimeError(f"reference restart failed on {name}:\n{r2.stdout[-2000:]}\n{r2.stderr[-2000:]}")
            for fld in FIELDS:
                fn = os.path.join(outdir, f"gas{fld}2.dat")
                if os.path.exists(fn):
                    data[f"restart_{fld}2"] = np.fromfile(fn, dtype=np.float64).reshape(nrad, nsec)
        # the first crash's dump (SourceEuler.c:276-284) and the crash marks
        for fld in ("dens", "vrad", "vtheta", "temper"):
            fn = os.path.join(outdir, f"gas{fld}999.dat")
            if os.path.exists(fn):
                data[f"{fld}999"] = np.fromfile(fn, dtype=np.float64).reshape(nrad, nsec)
        crash = re.findall(r"Crash! at time (\S+)", r.stdout)
        if crash:
            data["crash_time"] = np.array([float(x) for x in crash])
            # the per-substep marks ('c' crashed, '.' not; SourceEuler.c:285-287)
            # of a run without -d, where each DT's marks start a line
            rm = subprocess.run([REF, par], cwd=tmp, capture_output=True, text=True, timeout=3600)
            marks = "".join(m.group(1) for m in re.finditer(r"^([c.]+)(?:\t|$)", rm.stdout, re.M))
            data["crash_marks"] = np.frombuffer(marks.encode(), dtype=np.uint8)
        if stride is not None:
            data["row_stride"] = np.array(stride)
        tq = np.loadtxt(os.path.join(outdir, "tqwk0.dat"), ndmin=2)
        data["tqwk0"] = tq
        data["used_rad"] = np.loadtxt(os.path.join(outdir, "used_rad.dat"))
        # OutputElements rows (hash, time, a, e, i, Omega, omega, f, m, gas,
        # solid, r, x, y, z, vx, vy, vz; %.12g) of every NOUTELEMENTS-th DT
        orb = os.path.join(outdir, "nbody.orbits.dat")
        if os.path.exists(orb) and os.path.getsize(orb) > 0:
            data["nbody_orbits"] = np.loadtxt(orb, ndmin=2)
        # one ConditionCFL block per substep; SOR iterations per substep
        data["n_substeps"] = np.array(r.stdout.count("Timestep control information"))
        data["sor_niter"] = np.array([int(m) for m in re.findall(r"SOR: Niter = (\d+)", r.stdout)])
        data["omega_scan"] = np.array([[float(a), int(b)] for a, b in
                                       re.findall(r"SOR: omega = (\S+) -- Niter = (\d+)", r.stdout)])
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **data)
        print(f"{name}: {int(data['n_substeps'])} substeps, SOR niter {data['sor_niter'][:6]}..., "
              f"scan {len(data['omega_scan'])} solves")


def seeded_pebbles_of(disk, cfg):
    """The two-fluid workload's seeded pebble disk for a case's gas disk."""
    from fargo_thorin_amd.seed import seeded_pebbles
    return seeded_pebbles(disk, bulkdens=float(cfg.get("PebbleBulkDens", 1.0)),
                          adiabind=float(cfg.get("Adiabind", 1.4)))


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/thorin missing: make -f oracle/ref.mk")
    only = sys.argv[1:]
    for name, (cfg, planets, full) in CASES.items():
        if only and name not in only:
            continue
        run_case(name, cfg, planets, full)


if __name__ == "__main__":
    main()
