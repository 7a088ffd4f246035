"""An ensemble of disks advanced together (thorin_ensemble_*) against the same
disks advanced one by one with thorin_algogas: every disk's fields, step log
and FARGO shifts must be bit-identical, and the merged launches must carry
several disks each."""
import numpy as np
import pytest

from fargo_thorin_amd import Disk, Ensemble, configs, seeded_disk, seeded_pebbles
from fargo_thorin_amd._capi import host_nbody_fn
from tests.helpers import params_for, planets_for

pytestmark = pytest.mark.gpu

FIELDS = ["dens", "vrad", "vtheta", "energy", "temper", "soundspeed", "divv", "qplus", "opacity"]
PEB_FIELDS = ["pebbledens", "pebblevrad", "pebblevtheta", "pebblestokes"]


def _print_case(*a):
    print("[gpu-ensemble]", *a, flush=True)


def _members(sizes, n):
    cfgs = configs.ensemble_configs(n, sizes=sizes)
    out = []
    for c in cfgs:
        seed = c.pop("_seed")
        nrad, nsec = int(c["Nrad"]), int(c["Nsec"])
        disk = seeded_disk(nrad, nsec, sigma0=float(c["Sigma0"]), h=float(c["AspectRatio"]), seed=seed)
        out.append((c, disk))
    return out


def _make(cfg, disk, pebbles=False):
    d = Disk(params_for(cfg))
    d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    if pebbles:
        pb = seeded_pebbles(disk)
        d.init_pebbles(1, pb["pdens"], pb["pvrad"], pb["pvtheta"], pb["psize"])
    return d


def _log(lg):
    n = lg.nsub
    return (n, [lg.cfl_n[k] for k in range(n)], [lg.dt[k] for k in range(n)],
            [lg.sor_niter[k] for k in range(n)], [lg.sor_omega[k] for k in range(n)])


def _compare(a, b, fields):
    for f in fields:
        x, y = a.field(f), b.field(f)
        assert np.array_equal(x, y), f"{f}: max |diff| {np.max(np.abs(x - y)):.3e}"
    assert np.array_equal(a.ring_profile("nshift"), b.ring_profile("nshift"))


@pytest.mark.parametrize("sizes,n,steps,graphs", [([32, 48, 64], 4, 6, True), ([32, 48, 64], 4, 3, False),
                                                  ([512, 768], 3, 3, True), ([32, 48], 11, 3, True)])
def test_ensemble_matches_single_disks(sizes, n, steps, graphs):
    """Disks of uneven grids and parameters (alpha, Sigma0, h drawn as the
    ensemble workload draws them), from the first DT on (the SOR
    over-relaxation scan runs batched too); with graphs, the repeating
    substeps are replayed graphs, each replay checked against the disks
    stepped alone."""
    _print_case(f"sizes {sizes} n {n} steps {steps} graphs {graphs}")
    mem = _members(sizes, n)
    solo = [_make(c, dk) for c, dk in mem]
    ens_disks = [_make(c, dk) for c, dk in mem]
    ens = Ensemble(ens_disks)
    ens.set_graphs(graphs)
    nb = host_nbody_fn()
    ps = [planets_for(configs.PLANET_ZERO_MASS) for _ in mem]
    pe = [planets_for(configs.PLANET_ZERO_MASS) for _ in mem]
    for step in range(steps):
        if step == 1:
            l0, o0, _, _ = ens.stats()
        logs_s = [_log(d.algogas(p, nb)) for d, p in zip(solo, ps)]
        logs_e = [_log(lg) for lg in ens.algogas(pe, nb)]
        assert logs_s == logs_e, f"step {step}"
        for a, b in zip(ps, pe):
            assert (a.x[0], a.y[0], a.vx[0], a.vy[0]) == (b.x[0], b.y[0], b.vx[0], b.vy[0])
        for a, b in zip(solo, ens_disks):
            _compare(a, b, FIELDS)
    # after the first DT (each disk's over-relaxation scan: solves of
    # different lengths): the kernels of a substep are launched once for all
    # disks (each shared launch saves n-1 launches; all ops merged n-wide would
    # save ops*(n-1)/n); SOR iterations that only some disks need run for fewer
    l1, o1, cap, rep = ens.stats()
    ops = o1 - o0
    saved = ops - (l1 - l0)
    _print_case(f"launches {l1 - l0}, disk kernel operations {ops}, graphs captured {cap}, replays {rep}")
    assert ops >= (steps - 1) * n * 10, ops
    assert saved >= 0.8 * ops * (n - 1) / n, (l1 - l0, ops)
    if graphs:
        assert rep > 0
    else:
        assert cap == 0 and rep == 0
    ens.close()
    # the contexts stay usable on their own
    for d, p in zip(ens_disks, pe):
        d.algogas(p, nb)


def test_ensemble_two_fluid_and_gas_disks():
    """A two-fluid disk (pebble accretion onto a planet that feels the disk:
    the host round trips of the accretion and of the planet's disk forces)
    together with gas-only disks."""
    _print_case("two-fluid + gas")
    pcfg = configs.with_grid(configs.PEBBLES_1024x1536, 64, 96)
    pdisk = seeded_disk(64, 96, seed=777)
    mem = _members([48, 64], 2)
    solo = [_make(pcfg, pdisk, pebbles=True)] + [_make(c, dk) for c, dk in mem]
    ens_disks = [_make(pcfg, pdisk, pebbles=True)] + [_make(c, dk) for c, dk in mem]
    ens = Ensemble(ens_disks)
    nb = host_nbody_fn()
    ps = [planets_for(configs.PLANET_Q3E5, pcfg)] + [planets_for(configs.PLANET_ZERO_MASS) for _ in mem]
    pe = [planets_for(configs.PLANET_Q3E5, pcfg)] + [planets_for(configs.PLANET_ZERO_MASS) for _ in mem]
    for step in range(3):
        logs_s = [_log(d.algogas(p, nb)) for d, p in zip(solo, ps)]
        logs_e = [_log(lg) for lg in ens.algogas(pe, nb)]
        assert logs_s == logs_e, f"step {step}"
        assert (ps[0].mass[0], ps[0].solid[0], ps[0].vx[0]) == (pe[0].mass[0], pe[0].solid[0], pe[0].vx[0])
    _compare(solo[0], ens_disks[0], FIELDS + PEB_FIELDS)
    for a, b in zip(solo[1:], ens_disks[1:]):
        _compare(a, b, FIELDS)
    ens.close()


def test_launch_shapes_leave_results_unchanged():
    """Every launch shape the autotuner chooses from (SOR iterations per
    launch 1, 2 or 3, radial transport tiles of 8, 16 or 32 rings) gives the
    same bits, and the autotuner restores the state it times on."""
    _print_case("launch shapes")
    cfg, dk = _members([64], 1)[0]
    ref = _make(cfg, dk)
    nb = host_nbody_fn()
    pr = planets_for(configs.PLANET_ZERO_MASS)
    logs_r = [_log(ref.algogas(pr, nb)) for _ in range(4)]
    for nit in (1, 2, 3):
        for rows in (8, 16, 32):
            d = _make(cfg, dk)
            d.set_tuning(nit, rows)
            p = planets_for(configs.PLANET_ZERO_MASS)
            logs = [_log(d.algogas(p, nb)) for _ in range(4)]
            assert logs == logs_r, (nit, rows)
            _compare(ref, d, FIELDS)
            d.close()
    d = _make(cfg, dk)
    p = planets_for(configs.PLANET_ZERO_MASS)
    logs = [_log(d.algogas(p, nb))]
    t = d.autotune()                     # between steps
    _print_case(f"tuning {t}")
    assert t["sor_iters_per_launch"] in (1, 2, 3) and t["rad_rows"] in (8, 16, 32)
    assert min(t["sor_us_per_iter"]) > 0 and min(t["rad_us"]) > 0
    logs += [_log(d.algogas(p, nb)) for _ in range(3)]
    assert logs == logs_r
    _compare(ref, d, FIELDS)


def test_ensemble_disks_with_different_substep_counts():
    """A disk whose DT needs several CFL substeps next to disks that need one:
    the merged queues drift apart and rejoin (a disk that has finished its DT
    drops out of the merged launches), and every disk stays bit-identical."""
    _print_case("different substep counts")
    mem = _members([48, 64], 3)
    mem[1][0]["DT"] = "3.0"                        # several substeps per DT
    solo = [_make(c, dk) for c, dk in mem]
    ens_disks = [_make(c, dk) for c, dk in mem]
    ens = Ensemble(ens_disks)
    nb = host_nbody_fn()
    ps = [planets_for(configs.PLANET_ZERO_MASS) for _ in mem]
    pe = [planets_for(configs.PLANET_ZERO_MASS) for _ in mem]
    nsubs = set()
    for step in range(3):
        logs_s = [_log(d.algogas(p, nb)) for d, p in zip(solo, ps)]
        logs_e = [_log(lg) for lg in ens.algogas(pe, nb)]
        assert logs_s == logs_e, f"step {step}"
        nsubs |= {lg[0] for lg in logs_e}
    assert len(nsubs) > 1, nsubs
    for a, b in zip(solo, ens_disks):
        _compare(a, b, FIELDS)
    ens.close()
