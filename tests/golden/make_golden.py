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

# name: (config keys, planets, keep full fields?)
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
}


def run_case(name, cfg, planets, full):
    nrad, nsec = int(cfg["Nrad"]), int(cfg["Nsec"])
    with tempfile.TemporaryDirectory() as tmp:
        cfgdir = os.path.join(tmp, "in")
        outdir = os.path.join(tmp, "out")
        os.makedirs(cfgdir)
        os.makedirs(outdir)
        disk = seeded_disk(nrad, nsec, log_grid=cfg.get("RadialSpacing", "LOG")[0] in "lL")
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
        with open(par, "w") as f:
            f.write(configs.inpar_text(cfg, outdir + "/", pf, cfgdir))
        # -d: the reference prints its CFL block and "SOR: Niter = n" for every substep
        r = subprocess.run([REF, "-d", par], cwd=tmp, capture_output=True, text=True, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed on {name}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        ntot, ninterm = int(cfg["Ntot"]), int(cfg["Ninterm"])
        outs = list(range(0, ntot // ninterm + 1))
        data = {"nrad": nrad, "nsec": nsec, "ntot": ntot, "ninterm": ninterm, "outputs": np.array(outs)}
        for k in ("dens", "vrad", "vtheta", "temper"):
            data["init_" + k] = disk[k]
        for n in outs:
            for fld in FIELDS:
                fn = os.path.join(outdir, f"gas{fld}{n}.dat")
                if not os.path.exists(fn):
                    continue
                a = np.fromfile(fn, dtype=np.float64).reshape(nrad, nsec)
                data[f"sha_{fld}{n}"] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), dtype=np.uint8)
                data[f"ringmean_{fld}{n}"] = a.mean(axis=1)
                if full:
                    data[f"{fld}{n}"] = a
                else:
                    data[f"rows_{fld}{n}"] = a[:: max(1, nrad // 8)].copy()
        if name.startswith("pebbles"):
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
        tq = np.loadtxt(os.path.join(outdir, "tqwk0.dat"), ndmin=2)
        data["tqwk0"] = tq
        data["used_rad"] = np.loadtxt(os.path.join(outdir, "used_rad.dat"))
        # one ConditionCFL block per substep; SOR iterations per substep
        data["n_substeps"] = np.array(r.stdout.count("Timestep control information"))
        data["sor_niter"] = np.array([int(m) for m in re.findall(r"SOR: Niter = (\d+)", r.stdout)])
        data["omega_scan"] = np.array([[float(a), int(b)] for a, b in
                                       re.findall(r"SOR: omega = (\S+) -- Niter = (\d+)", r.stdout)])
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **data)
        print(f"{name}: {int(data['n_substeps'])} substeps, SOR niter {data['sor_niter'][:6]}..., "
              f"scan {len(data['omega_scan'])} solves")


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
