"""Parity of the HIP path (libthorin_amd.so on the GPU) with the oracle and
the reference.

Bars:
  * CFL substep count N, dt sequence and FARGO shifts Nshift: bit-exact;
  * SOR iterations per substep: equal;
  * fields: relative L-infinity <= 1e-10 (FP64; the only differences are
    the last-ulp differences of the GPU's exp/log/pow against glibc);
  * total mass with walls: conserved to round-off.
"""
import ctypes as C

import numpy as np
import pytest

from fargo_thorin_amd import Disk
from oracle import OracleDisk
from tests.helpers import FIELDS, Replay, case, golden, params_for, planets_for, rel_linf
from fargo_thorin_amd._capi import host_nbody_fn

pytestmark = pytest.mark.gpu

TOL = 1e-10
FULL_CASES = ["wall_lp85_64x128", "damp_bl94_48x96", "dampinit_zhu12_40x64", "planet_q3e-5_48x96", "irr_lp85_48x96",
              "damp_sm03_48x96", "gasacc_q1e-3_64x128", "flock_cavity_48x96", "dampflux_48x96", "drift_48x96",
              "cold_lp85_48x96", "std_nu_noind_48x96", "hill_q1e-4_48x96"]
# fields whose values are differences of nearly equal terms: compared
# relative to the field's own scale
STATE_FIELDS = ["dens", "vrad", "vtheta", "temper", "energy", "divv", "qplus", "opacity"]


def _print_case(*a):
    print("[gpu-parity]", *a, flush=True)


@pytest.mark.parametrize("name", FULL_CASES)
def test_gpu_vs_reference_golden(name):
    _print_case("golden", name)
    cfg, planets, _ = case(name)
    g = golden(name)
    d = Disk(params_for(cfg))
    d.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sysp = planets_for(planets, cfg)
    rep = Replay(g["tqwk0"])
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    niters, times = [], [d.state().physical_time]
    for f in STATE_FIELDS:
        if f"{f}0" in g:
            assert np.array_equal(d.field(f), g[f"{f}0"]), f"initial {f}"
    tq = g["tqwk0"]
    for i in range(ntot):
        lg = d.algogas(sysp, rep.cb)
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        times.append(d.state().physical_time)
        # torque log row of this DT (UpdateLog): time and position exact;
        # accelerations (block-ordered reductions) and torques to 1e-10 of
        # the acceleration scale
        assert lg.tq_rows >= 1
        row = [lg.tq[0][q] for q in range(7)]
        assert row[0] == tq[i, 0] and row[3] == tq[i, 3] and row[4] == tq[i, 4], f"DT {i}: torque row time/position"
        scale = abs(tq[i, 5]) + abs(tq[i, 6])
        for q in (1, 5, 6):
            assert abs(row[q] - tq[i, q]) <= 1e-10 * scale, f"DT {i}: torque column {q}"
        if tq[i, 1] != 0.0:
            assert abs(row[2] - tq[i, 2]) <= 1e-10 * scale * abs(tq[i, 2] / tq[i, 1]), f"DT {i}: normalised torque"
        if (i + 1) % ninterm == 0:
            n = (i + 1) // ninterm
            for f in STATE_FIELDS:
                key = f"{f}{n}"
                if key in g:
                    err = rel_linf(d.field(f), g[key])
                    assert err <= TOL, f"{key}: rel Linf {err:.3e}"
    assert niters == list(g["sor_niter"][:len(niters)])
    assert np.array_equal(np.array(times[:ntot]), g["tqwk0"][:ntot, 0])


@pytest.mark.parametrize("name,steps", [("wall_lp85_64x128", 25), ("damp_bl94_48x96", 25), ("irr_lp85_48x96", 25),
                                        ("damp_sm03_48x96", 25), ("gasacc_q1e-3_64x128", 25),
                                        ("flock_cavity_48x96", 25), ("dampflux_48x96", 25), ("drift_48x96", 25)])
def test_gpu_vs_oracle_stepwise(name, steps):
    """Same inputs and the same host N-body for both: per DT the CFL substep
    counts, dt, SOR iterations and Nshift agree exactly and the fields to 1e-10."""
    _print_case("stepwise", name, steps)
    cfg, planets, _ = case(name)
    g = golden(name)
    P = params_for(cfg)
    d = Disk(P)
    o = OracleDisk(P)
    for x in (d, o):
        x.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sd, so = planets_for(planets, cfg), planets_for(planets, cfg)
    nb = host_nbody_fn()
    for step in range(steps):
        lg = d.algogas(sd, nb)
        gn = [lg.cfl_n[k] for k in range(lg.nsub)]
        gdt = [lg.dt[k] for k in range(lg.nsub)]
        gni = [lg.sor_niter[k] for k in range(lg.nsub)]
        lo = o.algogas(so, nb)
        assert gn == [lo.cfl_n[k] for k in range(lo.nsub)], f"step {step}: CFL N"
        assert gdt == [lo.dt[k] for k in range(lo.nsub)], f"step {step}: dt"
        assert gni == [lo.sor_niter[k] for k in range(lo.nsub)], f"step {step}: SOR iterations"
        assert np.array_equal(d.ring_profile("nshift"), o.one_d("nshift")[: d.nrad]), f"step {step}: Nshift"
        # VMed is the ordered ring mean of vtheta (same summation order on
        # both sides); it inherits the fields' last-ulp differences
        vm_err = rel_linf(d.ring_profile("vmed"), o.one_d("vmed")[: d.nrad])
        assert vm_err <= 1e-13, f"step {step}: VMed {vm_err:.3e}"
        for f in STATE_FIELDS:
            err = rel_linf(d.field(f), o.field(f))
            assert err <= TOL, f"step {step} {f}: {err:.3e}"
        # planet masses (gas accretion: block-ordered sums on the GPU)
        for k in range(sd.nb):
            assert abs(sd.mass[k] - so.mass[k]) <= 1e-12 * abs(so.mass[k]), f"step {step}: planet {k} mass"
    assert d.state().physical_time == o.state().physical_time


def _tq_row(lg):
    return lg.tq_rows, [lg.tq[0][q] for q in range(7)]


@pytest.mark.parametrize("feel", [False, True])
def test_gpu_forces_enqueued_ahead_match_forces_enqueued_in_place(feel):
    """The next substep's planet forces are enqueued behind each substep's
    sync point, before the host has stepped the planets on, and used when
    the planet set then matches.  A disk whose own density is re-uploaded
    between DTs discards them (its forces are always enqueued in place); a
    disk whose planet is nudged between DTs no longer matches them.  Equal
    inputs give equal bits, step logs and torque rows either way."""
    name = "planet_q3e-5_48x96"
    _print_case("forces ahead", name, "feel", feel)
    cfg, planets, _ = case(name)
    g = golden(name)
    nb = host_nbody_fn()
    disks, systems = [], []
    for _ in range(4):
        d = Disk(params_for(cfg))
        d.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
        disks.append(d)
        systems.append(planets_for(planets, {"PlanetsFeelDisk": "YES"} if feel else None))
    for step in range(8):
        logs = [d.algogas(s, nb) for d, s in zip(disks, systems)]
        rows = [(_tq_row(lg), [lg.sor_niter[k] for k in range(lg.nsub)]) for lg in logs]
        assert rows[0] == rows[1] and rows[2] == rows[3], f"step {step}"
        disks[1].upload("dens", disks[1].field("dens"))
        disks[3].upload("dens", disks[3].field("dens"))
        for s in systems[2:]:
            s.x[0] *= 1.0 + 1e-7
        for a, b in ((disks[0], disks[1]), (disks[2], disks[3])):
            for f in STATE_FIELDS:
                assert np.array_equal(a.field(f), b.field(f)), f"step {step} {f}"
        for a, b in ((systems[0], systems[1]), (systems[2], systems[3])):
            assert (a.x[0], a.y[0], a.vx[0], a.vy[0]) == (b.x[0], b.y[0], b.vx[0], b.vy[0])
    assert not np.array_equal(disks[0].field("dens"), disks[2].field("dens"))


def test_gpu_mass_conservation_walls():
    """WALL boundaries and no damping: transport conserves the total mass."""
    _print_case("mass")
    cfg, planets, _ = case("wall_lp85_64x128")
    g = golden("wall_lp85_64x128")
    d = Disk(params_for(cfg))
    d.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    m0 = d.total_mass()
    sysp = planets_for(planets)
    for _ in range(20):
        d.algogas(sysp)
    m1 = d.total_mass()
    assert abs(m1 - m0) / m0 < 1e-13


def test_gpu_axisymmetric_disk_without_planet():
    """An axisymmetric disk without a planet: the GPU follows the oracle to
    1e-10, and the disk stays axisymmetric up to the small symmetry breaking
    the reference's own scheme has (red-black SOR stopped at a 1e-6 residual,
    round-off of the indirect term): below 1e-8 of each field's scale, with
    vrad measured against the orbital velocity."""
    _print_case("axisymmetric")
    cfg, _, _ = case("wall_lp85_64x128")
    from fargo_thorin_amd import Planets, seeded_disk
    disk = seeded_disk(64, 128, noise=0.0)
    P = params_for(cfg)
    d, o = Disk(P), OracleDisk(P)
    for x in (d, o):
        x.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    empty = Planets()
    empty.nb = 0
    nb = host_nbody_fn()
    for _ in range(5):
        d.algogas(empty, nb)
        o.algogas(empty, nb)
    vscale = np.max(np.abs(d.field("vtheta")))
    for f in ("dens", "vrad", "vtheta", "temper"):
        a = d.field(f)
        assert rel_linf(a, o.field(f)) <= TOL, f
        scale = vscale if f == "vrad" else np.max(np.abs(a))
        assert np.max(np.abs(a - a[:, :1])) / scale < 1e-8, f


@pytest.mark.parametrize("nrad,nsec,steps", [(256, 384, 5), (24, 1536, 3), (21, 1100, 3), (12, 2900, 2)])
def test_gpu_larger_grid_parity(nrad, nsec, steps):
    """Grids with several workgroups per ring, radial tiles that do not divide
    NRAD, and azimuthal rings of 1 to 6 cells per thread (1536 is the
    1024x1536 workload's ring; 1100 and 2900 leave threads without a cell):
    DTs against the oracle."""
    _print_case(f"{nrad}x{nsec}")
    from fargo_thorin_amd import configs, seeded_disk
    cfg = configs.with_grid(configs.DAMPED_1024x1536, nrad, nsec)
    disk = seeded_disk(nrad, nsec, seed=777)
    P = params_for(cfg)
    d, o = Disk(P), OracleDisk(P)
    for x in (d, o):
        x.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    sd, so = planets_for([(1e-13, 1.0)]), planets_for([(1e-13, 1.0)])
    nb = host_nbody_fn()
    for step in range(steps):
        lg = d.algogas(sd, nb)
        lo = o.algogas(so, nb)
        assert [lg.sor_niter[k] for k in range(lg.nsub)] == [lo.sor_niter[k] for k in range(lo.nsub)]
        assert np.array_equal(d.ring_profile("nshift"), o.one_d("nshift")[:nrad])
    for f in STATE_FIELDS:
        err = rel_linf(d.field(f), o.field(f))
        assert err <= TOL, f"{f}: {err:.3e}"


PEB_CASE = "pebbles_q3e-5_64x128"
PEB_FIELDS = ["pebbledens", "pebblevrad", "pebblevtheta"]


@pytest.mark.parametrize("name", [PEB_CASE, "pebheat_q3e-5_64x128"])
def test_gpu_pebbles_vs_reference_golden(name):
    """Two-fluid run of the reference (equilibrium pebble disk of a fresh run,
    back-reaction, particle diffusion, evaporation, a q = 3e-5 planet that
    feels the disk and accretes pebbles): the set-up is bit-exact, the SOR
    iterations and PhysicalTime equal, gas and pebble fields within 1e-10."""
    _print_case("pebbles golden", name)
    cfg, planets, _ = case(name)
    g = golden(name)
    d = Disk(params_for(cfg))
    d.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    d.init_pebbles(0)
    for f in PEB_FIELDS + ["pebblesize", "pebblestokes"]:
        assert np.array_equal(d.field(f), g[f"{f}0"]), f"initial {f}"
    sysp = planets_for(planets, cfg)
    m0 = sysp.mass[0]
    rep = Replay(g["tqwk0"], integrate=True)
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    niters, times = [], [d.state().physical_time]
    for i in range(ntot):
        lg = d.algogas(sysp, rep.cb)
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        times.append(d.state().physical_time)
        if (i + 1) % ninterm == 0:
            n = (i + 1) // ninterm
            for f in STATE_FIELDS + PEB_FIELDS:
                key = f"{f}{n}"
                if key in g:
                    err = rel_linf(d.field(f), g[key])
                    assert err <= TOL, f"{key}: rel Linf {err:.3e}"
    assert niters == list(g["sor_niter"][:len(niters)])
    assert np.array_equal(np.array(times[:ntot]), g["tqwk0"][:ntot, 0])
    assert sysp.mass[0] > m0


@pytest.mark.parametrize("nrad,nsec,steps", [(64, 128, 6), (256, 384, 4), (24, 1536, 3)])
def test_gpu_pebbles_vs_oracle_stepwise(nrad, nsec, steps):
    """The two-fluid workload's seeded pebble disk (Sigma_p = 0.01 Sigma,
    Stokes numbers 0.01-0.1 per ring) set up as a restart, same host N-body on
    both sides: CFL substeps, dt, SOR iterations and Nshift exact; gas and
    pebble fields and Stokes numbers within 1e-10; the planet's mass, solid
    mass and velocity after accretion within 1e-12 (the sums over the cells
    inside the accretion radius are tree sums on the GPU)."""
    _print_case(f"pebbles stepwise {nrad}x{nsec}")
    from fargo_thorin_amd import configs, seeded_disk, seeded_pebbles
    cfg = configs.with_grid(configs.PEBBLES_1024x1536, nrad, nsec)
    disk = seeded_disk(nrad, nsec, seed=777)
    pb = seeded_pebbles(disk)
    P = params_for(cfg)
    d, o = Disk(P), OracleDisk(P)
    for x in (d, o):
        x.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        x.init_pebbles(1, pb["pdens"], pb["pvrad"], pb["pvtheta"], pb["psize"])
    sd, so = planets_for(configs.PLANET_Q3E5, cfg), planets_for(configs.PLANET_Q3E5, cfg)
    nb = host_nbody_fn()
    for step in range(steps):
        lg = d.algogas(sd, nb)
        lo = o.algogas(so, nb)
        assert [lg.cfl_n[k] for k in range(lg.nsub)] == [lo.cfl_n[k] for k in range(lo.nsub)], f"step {step}: N"
        assert [lg.dt[k] for k in range(lg.nsub)] == [lo.dt[k] for k in range(lo.nsub)], f"step {step}: dt"
        assert [lg.sor_niter[k] for k in range(lg.nsub)] == [lo.sor_niter[k] for k in range(lo.nsub)]
        assert np.array_equal(d.ring_profile("nshift"), o.one_d("nshift")[:nrad]), f"step {step}: Nshift"
        for q in ("mass", "solid", "vx", "vy"):
            a, b = getattr(sd, q)[0], getattr(so, q)[0]
            assert abs(a - b) <= 1e-12 * abs(b), f"step {step}: planet {q} {a!r} vs {b!r}"
    for f in STATE_FIELDS + PEB_FIELDS:
        err = rel_linf(d.field(f), o.field(f))
        assert err <= TOL, f"{f}: {err:.3e}"
    # the library forms the Stokes numbers of the next substep at the end of
    # a substep (PebbleStokesNumbers, Pebbles.c:287-290, from the new density)
    from fargo_thorin_amd.seed import RHO2CGS
    tau = 2.0 * (1.0 / RHO2CGS) * o.field("pebblesize") / o.field("dens") * (1.0 / np.sqrt(1.4))
    assert rel_linf(d.field("pebblestokes"), tau) <= TOL
    # on the coarse 24-ring grid no cell centre lies inside Reff
    assert (sd.solid[0] > 0.0) == (so.solid[0] > 0.0) == (nrad > 24)


def test_gpu_powerlaw_opacity_vs_reference_lp85_on_cold_disk():
    """OpacityTable POWERLAW (kappa = k0 T^2) with k0 = 2e-4 on the cold disk
    (64-132 K) against the reference's LP85 run, whose ice-grain branch is
    2e-4 T^2 there: SOR iterations and time exact, fields to 1e-10."""
    name = "cold_lp85_48x96"
    _print_case("powerlaw", name)
    cfg, planets, _ = case(name)
    g = golden(name)
    d = Disk(params_for(cfg, OpacityTable="POWERLAW", PowerLawOpacity="2e-4"))
    d.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sysp = planets_for(planets, cfg)
    rep = Replay(g["tqwk0"])
    niters = []
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    for i in range(ntot):
        lg = d.algogas(sysp, rep.cb)
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        if (i + 1) % ninterm == 0:
            n = (i + 1) // ninterm
            for f in STATE_FIELDS:
                if f"{f}{n}" in g:
                    err = rel_linf(d.field(f), g[f"{f}{n}"])
                    assert err <= TOL, f"{f}{n}: {err:.3e}"
    assert niters == list(g["sor_niter"][:len(niters)])


def test_gpu_crash_run_and_dump_vs_reference():
    """The cold disk with one cell at -0.01 of its temperature: the crash
    marks of every substep ('cc..'), the first crash's time and its dump
    (gas{dens,vrad,vtheta,temper}999.dat) and the fields after three DTs
    against the reference; N and the marks against the oracle per DT."""
    name = "crash_cold_48x96"
    _print_case("crash", name)
    cfg, planets, _ = case(name)
    g = golden(name)
    P = params_for(cfg)
    d, o = Disk(P), OracleDisk(P)
    for x in (d, o):
        x.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sd, so = planets_for(planets, cfg), planets_for(planets, cfg)
    rep_d, rep_o = Replay(g["tqwk0"]), Replay(g["tqwk0"])
    marks, dumps, rcs = "", [], []
    for i in range(int(g["ntot"])):
        lg = d.algogas(sd, rep_d.cb)
        rcs.append(d.last_rc)
        gm = "".join("c" if lg.crashed[k] else "." for k in range(lg.nsub))
        gn = [lg.cfl_n[k] for k in range(lg.nsub)]
        if lg.crash_dumped:
            dumps.append(lg.crash_time)
        lo = o.algogas(so, rep_o.cb)
        assert gn == [lo.cfl_n[k] for k in range(lo.nsub)], f"DT {i}: N"
        assert gm == "".join("c" if lo.crashed[k] else "." for k in range(lo.nsub)), f"DT {i}: marks"
        marks += gm
    assert marks == bytes(g["crash_marks"]).decode()[: len(marks)]
    assert rcs[:2] == [1, 1] and rcs[2] == 0
    assert dumps == list(g["crash_time"])
    assert d.state().already_crashed == o.state().already_crashed == 2
    for f in ("dens", "vrad", "vtheta", "temper"):
        err = rel_linf(d.field("crash_" + f), g[f"{f}999"])
        assert err <= TOL, f"{f}999: {err:.3e}"
    for f in ("dens", "vrad", "vtheta", "temper", "energy"):
        err = rel_linf(d.field(f), g[f"{f}1"])
        assert err <= TOL, f"{f}1: {err:.3e}"


@pytest.mark.parametrize("nrad,nsec,wall", [(64, 128, False), (21, 1100, False), (37, 96, True), (256, 384, False),
                                            (12, 64, True)])
def test_gpu_fused_substep_chain_equals_separate_kernels(nrad, nsec, wall):
    """k_visc_chain (SubStep1, ViscousTerms, SubStep2 and the damping boundary
    around them in one kernel) against the separate kernels on the same disk:
    every field, the SOR iterations and the CFL counts bit for bit over
    several DTs.  The grids leave partial tiles in both directions; the walled
    disk has no damping and ImposeKeplerianEdges."""
    _print_case(f"fused chain {nrad}x{nsec} wall={wall}")
    from fargo_thorin_amd import configs, seeded_disk
    cfg = configs.with_grid(configs.DAMPED_1024x1536, nrad, nsec)
    if wall:
        cfg = dict(cfg, InnerBoundary="WALL")
    disk = seeded_disk(nrad, nsec, seed=4242)
    P = params_for(cfg)
    runs = []
    for fused in (1, 0):
        d = Disk(P)
        d.set_tuning(3, 8, fused_chain=fused)
        d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        sp = planets_for([(1e-5, 1.0)])
        nb = host_nbody_fn()
        logs = []
        for _ in range(4):
            lg = d.algogas(sp, nb)
            logs.append(([lg.cfl_n[k] for k in range(lg.nsub)], [lg.sor_niter[k] for k in range(lg.nsub)]))
        assert d.tuning()["fused_chain"] == fused
        runs.append((logs, {f: d.field(f, ghost=f in ("vrad", "vtheta")) for f in STATE_FIELDS + ["energyint"]}))
    assert runs[0][0] == runs[1][0]
    for f, a in runs[0][1].items():
        assert np.array_equal(a, runs[1][1][f]), f


@pytest.mark.parametrize("name", ["wall_lp85_64x128", "irr_lp85_48x96", "damp_sm03_48x96", "pebheat_q3e-5_64x128"])
def test_gpu_sor_assembly_equals_matrix_pass(name):
    """The energy system formed inside the SOR launches (the default) against
    the separate matrix pass that the SOR loads (sor_assembly 0), on the same
    disk: every field, the SOR iterations, the CFL counts and the planet mass
    bit for bit over 6 DTs.  The cases cover the stellar irradiation term of
    the right-hand side, the SM03 opacity, and a two-fluid run whose pebble
    accretion heating adds sources to it."""
    _print_case("sor assembly", name)
    cfg, planets, _ = case(name)
    g = golden(name)
    peb = name.startswith("pebheat")
    fields = STATE_FIELDS + (PEB_FIELDS if peb else [])
    runs = []
    for asm in (1, 0):
        d = Disk(params_for(cfg))
        d.set_tuning(3, 8, sor_assembly=asm)
        d.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
        if peb:
            d.init_pebbles(0)
        sysp = planets_for(planets, cfg)
        nb = host_nbody_fn()
        logs = []
        for _ in range(6):
            lg = d.algogas(sysp, nb)
            logs.append(([lg.cfl_n[k] for k in range(lg.nsub)], [lg.sor_niter[k] for k in range(lg.nsub)]))
        assert d.tuning()["sor_assembly"] == asm
        runs.append((logs, {f: d.field(f) for f in fields}, sysp.mass[0]))
        d.close()
    assert runs[0][0] == runs[1][0]
    for f, a in runs[0][1].items():
        assert np.array_equal(a, runs[1][1][f]), f
    assert runs[0][2] == runs[1][2]
