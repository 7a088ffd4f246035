"""The five workloads at their own sizes on the GPU.

  plain 128x256, 200 DTs     : against the reference's run (SOR iterations,
                               PhysicalTime, every 16th ring and all ring means
                               of outputs 1 and 2 to 1e-10; mass conserved
                               with walls to round-off)
  damped 1024x1536, 100 DTs  : against the reference's run (the same bars,
                               every 64th ring); its first DTs step by step
                               against the oracle: CFL N, dt, SOR iterations
                               and the FARGO shifts bit-exact
  two-fluid 1024x1536        : the seeded pebble disk through the reference's
                               own restart route, 100 of the 500 DTs: the
                               set-up bit-exact, SOR iterations and time
                               equal, gas and pebble rows to 1e-10, the
                               planet's mass and solid mass after every DT
                               against the reference's nbody.orbits.dat
  energy solve 2048x3072     : FP32 sweeps in FP64 refinement against FP64
                               SOR, both to the 1e-10 residual
  ensemble of 16 disks       : 512x768 to 1024x1536, two DTs, every disk bit
                               for bit equal to stepping it alone
"""
import numpy as np
import pytest

from fargo_thorin_amd import Disk, Ensemble, Params, configs, seeded_disk
from fargo_thorin_amd._capi import host_nbody_fn
from fargo_thorin_amd.seed import sor_disk
from oracle import OracleDisk
from tests.golden.make_golden import seeded_pebbles_of
from tests.helpers import Replay, case, golden, initial_disk, params_for, planets_for, rel_linf, sha

pytestmark = pytest.mark.gpu

TOL = 1e-10
GAS = ["dens", "vrad", "vtheta", "temper", "energy"]
PEB = ["pebbledens", "pebblevrad", "pebblevtheta"]


def _say(*a):
    print("[gpu-configs]", *a, flush=True)


def _check_output(d, g, n, fields):
    """Every stored ring (rows_*) to TOL relative L-inf, and every ring mean
    to TOL of the field's scale."""
    stride = int(g["row_stride"])
    for f in fields:
        if f"ringmean_{f}{n}" not in g:
            continue
        a = d.field(f)
        rows = g[f"rows_{f}{n}"]
        err = rel_linf(a[::stride], rows)
        assert err <= TOL, f"{f}{n} rows: rel Linf {err:.3e}"
        rm = g[f"ringmean_{f}{n}"]
        scale = max(np.max(np.abs(rows)), np.max(np.abs(rm)))
        merr = np.max(np.abs(a.mean(axis=1) - rm)) / scale
        assert merr <= TOL, f"{f}{n} ring means: {merr:.3e}"


def _golden_mass(d, g, n):
    surf = d.geometry()["surf"][: d.nrad]
    return float(np.sum(g[f"ringmean_dens{n}"] * d.nsec * surf))


def _run_against_golden(name, fields, pebble_setup=None, orbits=False):
    cfg, planets, _ = case(name)
    g = golden(name)
    disk = initial_disk(name)
    d = Disk(params_for(cfg))
    d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    if pebble_setup is not None:
        pebble_setup(d, g, disk, cfg)
    sysp = planets_for(planets, cfg)
    rep = Replay(g["tqwk0"], integrate=pebble_setup is not None)
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    # OutputElements writes a row after every DT (main.c:202-207), the
    # restarted run's last Ntot + 1 rows; the last is the extra DT
    orb = g["nbody_orbits"][-(ntot + 1):-1] if orbits else None
    niters, times, nsub = [], [d.state().physical_time], 0
    m0 = d.total_mass()
    for i in range(ntot):
        lg = d.algogas(sysp, rep.cb)
        nsub += lg.nsub
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        times.append(d.state().physical_time)
        if orb is not None:
            # OutputElements (%.12g): time, mass and solid mass times MassTaper
            # (of the DT's last substep: 0 in DT 0, 1 after)
            mt = d.state().mass_taper
            assert abs(orb[i, 1] - times[-1]) <= 1e-12 * times[-1], f"DT {i}: orbit row time"
            assert abs(sysp.mass[0] * mt - orb[i, 8]) <= 1e-11 * orb[i, 8], f"DT {i}: planet mass"
            assert abs(sysp.solid[0] * mt - orb[i, 10]) <= 1e-11 * orb[i, 10], f"DT {i}: solid mass"
        if (i + 1) % ninterm == 0:
            n = (i + 1) // ninterm
            _say(name, "output", n)
            _check_output(d, g, n, fields)
            gm = _golden_mass(d, g, n)
            assert abs(d.total_mass() - gm) <= 1e-12 * gm, f"output {n}: total mass"
    # the reference's main loop (main.c:171) runs DTs 0..Ntot, one more than
    # compared here; Ntot + 1 CFL blocks mean one substep in every DT
    if int(g["n_substeps"]) == ntot + 1:
        assert nsub == ntot
    else:
        assert nsub < int(g["n_substeps"])
    assert niters == list(g["sor_niter"][: len(niters)])
    assert np.array_equal(np.array(times[:ntot]), g["tqwk0"][:ntot, 0])
    return d, m0


def test_plain_128x256_200_steps_vs_reference():
    _say("plain 128x256 x 200")
    d, m0 = _run_against_golden("plain_128x256_200", GAS)
    # WALL at both edges, no damping: the transport conserves the mass
    assert abs(d.total_mass() - m0) <= 1e-13 * m0


def test_damped_1024x1536_100_steps_vs_reference():
    _say("damped 1024x1536 x 100")
    _run_against_golden("damped_1024x1536_100", GAS)


def test_damped_1024x1536_first_steps_vs_oracle():
    """Stepwise at the full grid: CFL N, dt, SOR iterations, Nshift and VMed
    as the oracle forms them; fields to 1e-10."""
    _say("damped 1024x1536 stepwise")
    name = "damped_1024x1536_100"
    cfg, planets, _ = case(name)
    disk = initial_disk(name)
    P = params_for(cfg)
    d, o = Disk(P), OracleDisk(P)
    for x in (d, o):
        x.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    sd, so = planets_for(planets), planets_for(planets)
    nb = host_nbody_fn()
    for step in range(3):
        lg = d.algogas(sd, nb)
        lo = o.algogas(so, nb)
        assert [lg.cfl_n[k] for k in range(lg.nsub)] == [lo.cfl_n[k] for k in range(lo.nsub)], f"step {step}: N"
        assert [lg.dt[k] for k in range(lg.nsub)] == [lo.dt[k] for k in range(lo.nsub)], f"step {step}: dt"
        assert [lg.sor_niter[k] for k in range(lg.nsub)] == [lo.sor_niter[k] for k in range(lo.nsub)]
        assert np.array_equal(d.ring_profile("nshift"), o.one_d("nshift")[: d.nrad]), f"step {step}: Nshift"
        assert rel_linf(d.ring_profile("vmed"), o.one_d("vmed")[: d.nrad]) <= 1e-13
    for f in GAS:
        err = rel_linf(d.field(f), o.field(f))
        assert err <= TOL, f"{f}: {err:.3e}"


def _seeded_pebbles_restart(d, g, disk, cfg):
    """The reference's route: a fresh run's output 0 (its InitializeFromFile
    state, bit for bit), read back as a restart with the seeded pebble files."""
    for k in ("dens", "vrad", "vtheta", "temper"):
        assert np.array_equal(sha(d.field(k)), g[f"fresh_sha_{k}0"]), f"fresh set-up {k}"
    st = {k: d.field(k) for k in ("dens", "vrad", "vtheta", "temper")}
    d.init(st["dens"], st["vrad"], st["vtheta"], st["temper"], mode=2)
    pb = seeded_pebbles_of(disk, cfg)
    d.init_pebbles(1, pb["pdens"], pb["pvrad"], pb["pvtheta"], pb["psize"])
    for k in ("dens", "vrad", "vtheta", "temper", "pebbledens", "pebblevrad", "pebblevtheta", "pebblesize"):
        assert np.array_equal(sha(d.field(k)), g[f"sha_{k}0"]), f"restart set-up {k}"


def test_two_fluid_1024x1536_100_steps_vs_reference():
    _say("two-fluid 1024x1536 x 100")
    d, _ = _run_against_golden("pebbles_1024x1536_100", GAS + PEB, pebble_setup=_seeded_pebbles_restart,
                               orbits=True)


def test_energy_solve_2048x3072_mixed_precision_vs_fp64_sor():
    _say("energy solve 2048x3072")
    cfg = configs.SOR_2048x3072
    d = Disk(Params(**cfg, OutputDir="/nonexistent-thorin-out/"))
    disk = sor_disk(2048, 3072)
    d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    dt = d.rad_system(0.0)
    assert dt > 0
    p = d.rad_solve(0, tol=1e-10, omega=1.8)
    Tp = d.field("temper")
    m = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=16)
    Tm = d.field("temper")
    _say(f"fp64 sor {p['ms']:.2f} ms {p['sweeps']} sweeps; fp32 refine {m['ms']:.2f} ms {m['sweeps']} sweeps")
    assert p["converged"] and m["converged"]
    assert m["resid"] <= 1e-10 * m["resid0"]
    assert p["resid"] <= 1e-9 * p["resid0"]
    assert np.max(np.abs(Tm - Tp)) / np.max(np.abs(Tp)) < 1e-9
    assert m["ms"] < p["ms"]


def test_ensemble_16_disks_full_sizes_match_single_disks():
    _say("ensemble 16 disks 512x768-1024x1536")
    cfgs = configs.ensemble_configs(16)
    solo, ens_disks = [], []
    for c in cfgs:
        c = dict(c)
        seed = c.pop("_seed")
        nrad, nsec = int(c["Nrad"]), int(c["Nsec"])
        disk = seeded_disk(nrad, nsec, sigma0=float(c["Sigma0"]), h=float(c["AspectRatio"]), seed=seed)
        for lst in (solo, ens_disks):
            x = Disk(params_for(c))
            x.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
            lst.append(x)
    ens = Ensemble(ens_disks)
    nb = host_nbody_fn()
    ps = [planets_for(configs.PLANET_ZERO_MASS) for _ in cfgs]
    pe = [planets_for(configs.PLANET_ZERO_MASS) for _ in cfgs]
    for step in range(2):
        ls = [d.algogas(p, nb) for d, p in zip(solo, ps)]
        ls = [(lg.nsub, [lg.sor_niter[k] for k in range(lg.nsub)]) for lg in ls]
        le = [(lg.nsub, [lg.sor_niter[k] for k in range(lg.nsub)]) for lg in ens.algogas(pe, nb)]
        assert ls == le, f"step {step}"
    for k, (a, b) in enumerate(zip(solo, ens_disks)):
        for f in ("dens", "vrad", "vtheta", "energy"):
            assert np.array_equal(a.field(f), b.field(f)), f"disk {k} {f}"
        assert np.array_equal(a.ring_profile("nshift"), b.ring_profile("nshift")), f"disk {k} nshift"
    launches, ops, _, _ = ens.stats()
    assert ops > launches
