#!/usr/bin/env python3
"""Throughput of the GPU hydro step (AlgoGas) at 1024x1536 FP64 on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

A step is one DT of AlgoGas (CFL substeps, forces, source terms, viscosity,
implicit radiative diffusion with SOR, FARGO transport) on the seeded
1024x1536 disk with damping zones at both edges; the fields stay resident
in HBM.  The first warmup step includes the reference's one-off search of
the SOR relaxation parameter and is never timed.

Printed: one JSON line with steps/s, the HIP-event roofline of the dominant
kernel (the one moving the most algorithmic bytes per step, timed in a
region of its own after the throughput region), the whole step's roofline,
the reference's own CPU builds (oracle/_ref: sequential, OpenMP, MPI) timed
on this host on a few DTs of the same workload, and the studies (energy
solve at 2048x3072, two-fluid step, 16-disk ensemble).
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s @1024×1536 fp64; SOR ms/step CNN-warm+FP32-refine vs plain @1e-10"
HBM_PEAK_GBS = 8000.0

# Compulsory FP64 traffic per cell and launch of each kernel: 8 B per full
# field the kernel must read or write once.  SOR: 6 matrix coefficients + T
# read, T written = 64 B per cell and iteration over rings 1..NRAD-2; its
# first launch of a solve also assembles the system (k_matrix's 104 B/cell:
# rho, cs, D, div v, Q-, Q+, T read, five diagonals and the RHS written).
ALG_BYTES_PER_CELL = {
    "k_sor_iter": 64, "k_transport_theta": 104, "k_transport_rad": 80, "k_forces": 48,
    "k_substep1": 64, "k_divstress": 56, "k_visc_update": 64, "k_substep2": 56, "k_matrix": 104, "k_velocities": 56,
    "k_indirect": 48, "k_eos_cfl": 48, "k_eos_cfl_vel": 80,
    "k_substep3": 96, "k_visc_chain": 72,
}
# Whole step: the fused-dataflow minimum of the hydro step, 82 field passes
# per substep (656 B/cell) plus 64 B/cell per SOR iteration.
STEP_BYTES_PER_CELL = 656
SOR_BYTES_PER_CELL_ITER = 64
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (KiB per dispatch), scaled by what the counters read for a 1 GiB device copy
# (FETCH_SIZE reads half the bytes moved, WRITE_SIZE all of them).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_damped_1024x1536.txt")
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0


def measured_traffic(kname, nit=None):
    """Bytes per launch of kname from the committed counter summary, or None.
    For k_sor_iter the row of the launch shape with nit iterations."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("#"):
                    continue
                name = line[:35].strip()
                parts = line[35:].split()
                if len(parts) < 2 or name.split("<")[0] != kname:
                    continue
                if nit is not None and (not name.startswith(f"{kname}<{nit},") or name.endswith("true>")):
                    continue            # (the launches that iterate, not the forming first one)
                return FETCH_SCALE * float(parts[0]) * 1024 + WRITE_SCALE * float(parts[1]) * 1024
    except OSError:
        return None
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nrad", type=int, default=1024)
    ap.add_argument("--nsec", type=int, default=1536)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sor-study", action="store_true")
    ap.add_argument("--no-two-fluid", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-cnn", action="store_true",
                    help="ensemble study without its warm-start part (the disks' stepping alone)")
    ap.add_argument("--study", choices=["sor", "two-fluid", "ensemble"], default=None,
                    help="run only this study and print its JSON (for profiling)")
    ap.add_argument("--roofline-kernel", default="auto",
                    help="kernel of the roofline line (auto: the most algorithmic bytes per step)")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="only time a 1 GiB device copy with HIP events and print its GB/s")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (3 times), the reference for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 1 << 27                       # 2^27 doubles = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))


def copy_bandwidth():
    """Device-to-device copy of 1 GiB (read + write = 2 GiB per copy), timed
    with events: the HBM bandwidth a plain streaming kernel reaches here."""
    import torch
    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(5):
        y.copy_(x)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 5
    return {"bytes_per_copy": 2 * n * 8, "ms": ms, "GBps": 2 * n * 8 / (ms * 1e-3) / 1e9}


CPU_CORES = 16      # the host CPU share of one GPU on the benchmark box


def _ref_wall(cmd, cfgdir, pf, tmp, cfg, env, tag):
    out = os.path.join(tmp, tag)
    os.makedirs(out)
    par = os.path.join(tmp, f"{tag}.par")
    with open(par, "w") as f:
        from fargo_thorin_amd import configs
        f.write(configs.inpar_text(cfg, out + "/", pf, cfgdir))
    t0 = time.perf_counter()
    r = subprocess.run(cmd + [par], cwd=tmp, capture_output=True, text=True, timeout=900, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"{tag}: {(r.stderr or r.stdout)[-300:]}")
    return time.perf_counter() - t0


def cpu_baseline_reference(nrad, nsec):
    """The reference's own builds (oracle/_ref, compiled from its sources by
    oracle/ref.mk) timed on this host: wall(Ntot = 1 + n) - wall(Ntot = 1)
    over n DTs (AlgoGas runs Ntot + 1 times; the set-up and the first DT's
    over-relaxation scan cancel):
      * sequential (1 core, mpi_dummy), n = 5;
      * OpenMP (make_omp.sh: -fopenmp, CPU_CORES threads), n = 10;
      * MPI (BUILD=parallel, mpiexec -n CPU_CORES), n = 10.
    The value reported is the fastest parallel build; the 128x256 workload's
    200-DT rate of the sequential and OpenMP builds is added."""
    from fargo_thorin_amd import configs
    from fargo_thorin_amd.seed import seeded_disk, write_cfg
    refdir = os.path.join(ROOT, "oracle", "_ref")
    seq = os.path.join(refdir, "thorin")
    if not os.path.exists(seq):
        return None
    tmp = tempfile.mkdtemp(prefix="thorin_cpu_")
    pf = os.path.join(tmp, "planet.cfg")
    with open(pf, "w") as f:
        f.write(configs.planet_text(configs.PLANET_ZERO_MASS))
    env1 = dict(os.environ, OMP_NUM_THREADS="1")
    envn = dict(os.environ, OMP_NUM_THREADS=str(CPU_CORES))
    builds = {"sequential": ([seq], env1, 1, 5)}
    if os.path.exists(os.path.join(refdir, "thorin_omp")):
        builds["openmp"] = ([os.path.join(refdir, "thorin_omp")], envn, CPU_CORES, 10)
    mpiexec = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.path.exists("/opt/conda/bin/mpiexec") else None)
    if mpiexec and os.path.exists(os.path.join(refdir, "thorin_mpi")):
        builds["mpi"] = ([mpiexec, "-n", str(CPU_CORES), os.path.join(refdir, "thorin_mpi")], env1, CPU_CORES, 10)
    res = {}
    for grid, cases in (((nrad, nsec), builds),
                        ((128, 256), {k: v for k, v in builds.items() if k in ("sequential", "openmp")})):
        gr, gs = grid
        cfgdir = os.path.join(tmp, f"in{gr}x{gs}")
        os.makedirs(cfgdir)
        disk = seeded_disk(gr, gs)
        for k in ("dens", "vrad", "vtheta", "temper"):
            write_cfg(os.path.join(cfgdir, f"gas{k}.cfg"), disk[k])
        base = configs.DAMPED_1024x1536 if grid == (nrad, nsec) else configs.PLAIN_128x256
        for name, (cmd, env, cores, n) in cases.items():
            if grid == (128, 256):
                n = 200
            print(f"[bench] cpu baseline: reference {name} build, {gr}x{gs}, {n} DTs", file=sys.stderr, flush=True)
            try:
                w = {}
                for ntot in (1, 1 + n):
                    cfg = configs.with_grid(base, gr, gs, Ntot=ntot, Ninterm=ntot)
                    w[ntot] = _ref_wall(cmd, cfgdir, pf, tmp, cfg, env, f"{name}_{gr}x{gs}_{ntot}")
                per = (w[1 + n] - w[1]) / n
                res[f"{name}_{gr}x{gs}"] = {"steps_per_s": 1.0 / per, "cores": cores, "dts": n,
                                            "wall_s": [w[1], w[1 + n]]}
            except Exception as e:  # one build failing leaves the others
                res[f"{name}_{gr}x{gs}"] = {"error": str(e)}
    shutil.rmtree(tmp, ignore_errors=True)
    par_keys = [k for k in (f"openmp_{nrad}x{nsec}", f"mpi_{nrad}x{nsec}") if "steps_per_s" in res.get(k, {})]
    best = max(par_keys, key=lambda k: res[k]["steps_per_s"]) if par_keys else f"sequential_{nrad}x{nsec}"
    if "steps_per_s" not in res.get(best, {}):
        return {"error": "no reference build ran", "builds": res}
    b = res[best]
    return {"value": b["steps_per_s"], "unit": "steps/s", "cores": b["cores"], "kind": "reference",
            "sample": f"{b['dts']} DTs of the {nrad}x{nsec} damped workload, reference {best.split('_')[0]} build "
                      f"(oracle/_ref), {b['cores']} cores: wall(Ntot={1 + b['dts']}) - wall(Ntot=1)",
            "builds": res}


def train_for(net, seconds, batch):
    """Adam on the network's pairs for about `seconds`, the learning rate
    stepped down over the run (1e-3, 3e-4, 1e-4, 3e-5 at 30, 60 and 85% of
    it): (iterations, loss of the first batch's first pass, loss of the last
    batch's last pass)"""
    t0 = time.perf_counter()
    iters, first, last = 0, None, 0.0
    while time.perf_counter() - t0 < seconds:
        fr = (time.perf_counter() - t0) / seconds
        lr = 1e-3 if fr < 0.3 else 3e-4 if fr < 0.6 else 1e-4 if fr < 0.85 else 3e-5
        lf, last = net.train(batch, lr)
        first = lf if first is None else first
        iters += batch
    return iters, first, last


def sor_study(nrad=2048, nsec=3072, train_s=25.0, coarsen=2):
    """The energy equation's implicit solve at nrad x nsec (BL94 opacity,
    T 20-2000 K, dt = the hydro CFL step), all to a 1e-10 relative FP64
    residual at omega = 1.8:
      * plain FP64 red-black SOR (the reference's solver);
      * FP32 red-black correction sweeps in FP64 iterative refinement, four
        sweeps per launch on LDS tiles; sweeps per round picked by timing 32
        and 64 once;
      * the same from the learned warm start: the U-Net on the grid
        coarsened `coarsen` times, trained here on the GPU (Adam with a
        stepped learning rate, residual loss of the interpolated guess) for
        about train_s seconds on 12 pairs the solver generates from other
        seeds and steps (0.5, 1, 2 x CFL dt); the study disk itself is not
        among them.  Its time includes the network's inference and the
        residual of its guess.
    Errors are stated against the FP64 SOR result."""
    import numpy as np

    from fargo_thorin_amd import Disk, Params, WarmStart, configs
    from fargo_thorin_amd.seed import sor_disk
    cfg = configs.with_grid(configs.SOR_2048x3072, nrad, nsec)
    P = Params(**cfg, OutputDir="/nonexistent-thorin-out/")
    d = Disk(P)

    def assemble(seed, mult=1.0):
        disk = sor_disk(nrad, nsec, seed=seed)
        d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        dt0 = d.rad_system(0.0)
        return d.rad_system(dt0 * mult) if mult != 1.0 else dt0

    tol, omega = 1e-10, 1.8
    # training pairs (other disks) and the network
    t0 = time.perf_counter()
    net = WarmStart(seed=3)
    net.set_coarsen(coarsen)
    net.set_loss(1)
    for seed in range(4):
        for mult in (0.5, 1.0, 2.0):
            assemble(200 + seed, mult)
            d.rad_solve(1, tol=1e-11, omega=omega, inner_sweeps=64)
            net.add_pair(d)
    t_pairs = time.perf_counter() - t0
    t1 = time.perf_counter()
    iters, *loss = train_for(net, train_s, 48)
    t_train = time.perf_counter() - t1
    # the study disk
    dt = assemble(12345)
    tune = {}
    for inner in (32, 64):
        tune[inner] = d.rad_solve(1, tol=tol, omega=omega, inner_sweeps=inner)["ms"]
    inner = min(tune, key=tune.get)
    plain = min((d.rad_solve(0, tol=tol, omega=omega) for _ in range(2)), key=lambda r: r["ms"])
    t_plain = d.field("temper")
    mixed = min((d.rad_solve(1, tol=tol, omega=omega, inner_sweeps=inner) for _ in range(2)), key=lambda r: r["ms"])
    t_mixed = d.field("temper")
    warm = min((d.rad_solve(1, tol=tol, omega=omega, inner_sweeps=inner, warm=net) for _ in range(2)),
               key=lambda r: r["ms"])
    t_warm = d.field("temper")
    err = lambda t: float(np.max(np.abs(t - t_plain)) / np.max(np.abs(t_plain)))
    out = {"workload": f"sor_{nrad}x{nsec}", "dt": dt, "tol": tol, "omega": omega,
           "plain_fp64": {"ms": plain["ms"], "sweeps": plain["sweeps"],
                          "rel_resid": plain["resid"] / plain["resid0"]},
           "fp32_refine": {"ms": mixed["ms"], "sweeps": mixed["sweeps"], "rounds": mixed["rounds"],
                           "inner_sweeps": inner, "sweeps_per_launch": 4,
                           "rel_resid": mixed["resid"] / mixed["resid0"], "rel_err_vs_fp64_sor": err(t_mixed)},
           "cnn_warm_fp32_refine": {"ms": warm["ms"], "sweeps": warm["sweeps"], "rounds": warm["rounds"],
                                    "warm_used": warm["warm_used"],
                                    "guess_rel_resid": warm["resid_warm"] / warm["resid0"],
                                    "rel_resid": warm["resid"] / warm["resid0"], "rel_err_vs_fp64_sor": err(t_warm),
                                    "coarsen": coarsen,
                                    "training": {"pairs": net.n_pairs(), "iters": iters, "seconds": t_train,
                                                 "pair_generation_s": t_pairs, "loss_first": loss[0],
                                                 "loss_last": loss[1], "loss": "residual of the guess"}},
           "speedup_fp32_refine": plain["ms"] / mixed["ms"],
           "speedup_cnn_warm_fp32_refine": plain["ms"] / warm["ms"],
           "inner_sweep_timing_ms": tune}
    net.close()
    d.close()
    return out


def two_fluid_study(nrad=1024, nsec=1536, steps=20, warmup=3):
    """The two-fluid workload: the damped disk plus the seeded pebble disk
    (Sigma_p = 0.01 Sigma, Stokes numbers 0.01-0.1 per ring), back-reaction,
    particle diffusion, evaporation and a q = 3e-5 planet that feels the disk
    and accretes pebbles (one host sync per substep for the planet update).
    Steps/s over `steps` DTs after `warmup`, and the per-kernel split."""
    from fargo_thorin_amd import Disk, Params, configs, seeded_disk, seeded_pebbles
    from fargo_thorin_amd import make_planets
    cfg = configs.with_grid(configs.PEBBLES_1024x1536, nrad, nsec)
    P = Params(**cfg, OutputDir="/nonexistent-thorin-out/")
    disk = seeded_disk(nrad, nsec)
    pb = seeded_pebbles(disk)
    d = Disk(P)
    d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    d.init_pebbles(1, pb["pdens"], pb["pvrad"], pb["pvtheta"], pb["psize"])
    sysp = make_planets([3e-5], [1.0], feel_disk=True)
    for _ in range(warmup):
        d.algogas(sysp)
    d.sync()
    t0 = time.perf_counter()
    substeps = 0
    for _ in range(steps):
        substeps += d.algogas(sysp).nsub
    d.sync()
    el = time.perf_counter() - t0
    d.profile()
    for _ in range(3):
        d.algogas(sysp)
    prof = d.profile_read()
    d.profile_off()
    m_acc = sysp.solid[0]
    d.close()
    return {"workload": f"pebbles_{nrad}x{nsec}", "steps": steps, "substeps": substeps,
            "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "accreted_solid_mass": m_acc,
            "kernel_ms_per_step": {k: v[0] / 3 for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}}


def ensemble_study(n=16, steps=10, warmup=2, cnn=True):
    """The ensemble workload: 16 damped disks of uneven grids (512x768 to
    1024x1536) with seeded alpha, Sigma0 and h, advanced one DT at a time
    (a) disk after disk with thorin_algogas and (b) together with
    thorin_ensemble_algogas, where each kernel of a substep is one launch over
    all disks.  Disk-steps/s of both over `steps` DTs after `warmup` (the
    first DT holds each disk's SOR over-relaxation scan), launches per DT, and
    whether both runs end with bit-identical fields."""
    import numpy as np

    from fargo_thorin_amd import Disk, Ensemble, Params, configs, make_planets, seeded_disk
    mem = []
    for c in configs.ensemble_configs(n):
        c = dict(c)
        seed = c.pop("_seed")
        nr, ns = int(c["Nrad"]), int(c["Nsec"])
        mem.append((c, seeded_disk(nr, ns, sigma0=float(c["Sigma0"]), h=float(c["AspectRatio"]), seed=seed)))

    def make():
        ds = []
        for c, dk in mem:
            d = Disk(Params(**c, OutputDir="/nonexistent-thorin-out/"))
            d.init(dk["dens"], dk["vrad"], dk["vtheta"], dk["temper"])
            ds.append(d)
        return ds

    cells = sum(int(c["Nrad"]) * int(c["Nsec"]) for c, _ in mem)
    out = {"workload": f"ensemble_{n}", "disks": n, "cells": cells,
           "grids": [f"{c['Nrad']}x{c['Nsec']}" for c, _ in mem], "steps": steps, "warmup": warmup}
    # (a) one disk after the other
    A = make()
    pa = [make_planets([1e-13], [1.0]) for _ in mem]
    for _ in range(warmup):
        for d, p in zip(A, pa):
            d.algogas(p)
    for d in A:
        d.sync()
    t0 = time.perf_counter()
    subs = 0
    for _ in range(steps):
        for d, p in zip(A, pa):
            subs += d.algogas(p).nsub
    for d in A:
        d.sync()
    el_a = time.perf_counter() - t0
    out["one_by_one"] = {"disk_steps_per_s": n * steps / el_a, "ms_per_dt": 1e3 * el_a / steps,
                         "substeps": subs}
    # (b) the ensemble, merged substeps replayed as graphs; (c) without graphs
    same = True
    for key, graphs in (("ensemble", True), ("ensemble_no_graphs", False)):
        B = make()
        pb = [make_planets([1e-13], [1.0]) for _ in mem]
        E = Ensemble(B)
        E.set_graphs(graphs)
        E.algogas(pb)
        out["tuning"] = E.autotune()
        for _ in range(warmup - 1):
            E.algogas(pb)
        l0, o0, c0, r0 = E.stats()
        t0 = time.perf_counter()
        subs = 0
        for _ in range(steps):
            subs += sum(lg.nsub for lg in E.algogas(pb))
        el_b = time.perf_counter() - t0
        l1, o1, c1, r1 = E.stats()
        out[key] = {"disk_steps_per_s": n * steps / el_b, "ms_per_dt": 1e3 * el_b / steps, "substeps": subs,
                    "launches_per_dt": (l1 - l0) / steps, "disk_kernel_ops_per_dt": (o1 - o0) / steps,
                    "graph_captures": c1 - c0, "graph_replays": r1 - r0}
        for a, b in zip(A, B):
            for f in ("dens", "vrad", "vtheta", "energy"):
                same &= bool(np.array_equal(a.field(f), b.field(f)))
        if graphs and cnn:
            out["cnn_warm_start"] = ensemble_cnn_study(E, B, pb)
        E.close()
        for d in B:
            d.close()
    out["speedup"] = out["one_by_one"]["ms_per_dt"] / out["ensemble"]["ms_per_dt"]
    out["bit_identical"] = same
    for d in A:
        d.close()
    return out


def ensemble_cnn_study(E, disks, planets, train_s=20.0, tol=1e-10, omega=1.8, inner=16, coarsen=2):
    """The learned warm start on the ensemble: the network is trained on the
    GPU on pairs that the ensemble's own disks give (the first half of the
    disks, at two times of the run: each pair is the system of the disk's
    substep and its solution to 1e-11), then the energy solve of every disk
    of the other half is timed to a 1e-10 relative residual three ways: the
    reference's FP64 SOR, FP32 sweeps in FP64 refinement, and the same from
    the network's guess (the network on the grid coarsened `coarsen`
    times).  The solves use the disk's own substep dt."""
    from fargo_thorin_amd import WarmStart
    n = len(disks)
    train, evald = list(range(n // 2)), list(range(n // 2, n))

    def sub_dt(k):
        lg = E.logs[k]
        return lg.dt[0] if lg.nsub else 1e-3
    net = WarmStart(seed=5)
    net.set_coarsen(coarsen)
    net.set_loss(1)
    t0 = time.perf_counter()
    for rnd in range(2):
        if rnd:
            for _ in range(5):
                E.algogas(planets)
        for k in train:
            d = disks[k]
            d.rad_system(sub_dt(k))
            d.rad_solve(1, tol=1e-11, omega=omega, inner_sweeps=64)
            net.add_pair(d)
    t_pairs = time.perf_counter() - t0
    t1 = time.perf_counter()
    iters, *loss = train_for(net, train_s, 32)
    t_train = time.perf_counter() - t1
    rows = []
    for k in evald:
        d = disks[k]
        dt = sub_dt(k)
        res = {}
        for key, kw in (("plain_fp64", dict(method=0)), ("fp32_refine", dict(method=1, inner_sweeps=inner)),
                        ("cnn_warm_fp32_refine", dict(method=1, inner_sweeps=inner, warm=net))):
            best = None
            for _ in range(2):
                d.rad_system(dt)
                r = d.rad_solve(tol=tol, omega=omega, **kw)
                best = r if best is None or r["ms"] < best["ms"] else best
            res[key] = best
        rows.append({"grid": f"{d.nrad}x{d.nsec}", "dt": dt,
                     "ms": {k: v["ms"] for k, v in res.items()},
                     "sweeps": {k: v["sweeps"] for k, v in res.items()},
                     "warm_used": res["cnn_warm_fp32_refine"]["warm_used"],
                     "guess_rel_resid": res["cnn_warm_fp32_refine"]["resid_warm"] /
                     res["cnn_warm_fp32_refine"]["resid0"]})
    tot = {k: sum(r["ms"][k] for r in rows) for k in rows[0]["ms"]}
    net_pairs = net.n_pairs()
    net.close()
    return {"train_disks": len(train), "eval_disks": len(evald), "pairs": net_pairs, "pair_generation_s": t_pairs,
            "train_iters": iters, "train_s": t_train, "loss_first": loss[0], "loss_last": loss[1],
            "tol": tol, "omega": omega, "inner_sweeps": inner, "coarsen": coarsen,
            "ms_per_disk_step": {k: v / len(rows) for k, v in tot.items()},
            "speedup_vs_plain": {k: tot["plain_fp64"] / v for k, v in tot.items()},
            "disks": rows}


def c_driver_timing(nrad=1024, nsec=1536, ntot=60):
    """The same workload stepped by the library's C driver (thorin_amd -t):
    the AlgoGas calls timed in C, with no Python between them.  The driver
    starts from the analytic InitEuler disk (InitializeFromFile NO), writes
    outputs only at the first and last DT, and no torque files."""
    import subprocess
    import tempfile

    from fargo_thorin_amd import configs
    exe = os.path.join(ROOT, "fargo_thorin_amd", "thorin_amd")
    cfg = configs.with_grid(configs.DAMPED_1024x1536, nrad, nsec, Ntot=ntot, Ninterm=ntot,
                            InitializeFromFile="NO", WriteTorqueFiles="NO")
    with tempfile.TemporaryDirectory() as tmp:
        pf = os.path.join(tmp, "planet.cfg")
        with open(pf, "w") as f:
            f.write(configs.planet_text(configs.PLANET_ZERO_MASS))
        par = os.path.join(tmp, "in.par")
        with open(par, "w") as f:
            f.write(configs.inpar_text(cfg, tmp + "/out/", pf, tmp))
        r = subprocess.run([exe, "-t", par], capture_output=True, text=True, timeout=600)
    line = [x for x in r.stdout.splitlines() if x.startswith("timing:")]
    if r.returncode != 0 or not line:
        return {"error": (r.stderr or r.stdout)[-400:]}
    w = line[-1].split()
    return {"calls": int(w[1]), "ms_per_step": float(w[6]), "steps_per_s": 1e3 / float(w[6]),
            "disk": "analytic InitEuler profiles (no noise)"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.copy_bandwidth:
        print(json.dumps(copy_bandwidth()))
        return
    if args.study == "sor":
        print(json.dumps(sor_study()))
        return
    if args.study == "two-fluid":
        print(json.dumps(two_fluid_study()))
        return
    if args.study == "ensemble":
        print(json.dumps(ensemble_study(cnn=not args.no_cnn)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 (radial slabs over several "
                                   "GPUs are a design, not built)"}))
        sys.exit(2)
    import numpy as np

    from fargo_thorin_amd import Disk, Params, configs, make_planets, seeded_disk

    nrad, nsec = args.nrad, args.nsec
    cfg = configs.with_grid(configs.DAMPED_1024x1536, nrad, nsec)
    P = Params(**cfg, OutputDir="/nonexistent-thorin-out/")
    disk = seeded_disk(nrad, nsec)
    d = Disk(P)
    d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    sysp = make_planets([1e-13], [1.0])
    print(f"[bench] {nrad}x{nsec}: warmup {args.warmup} (first DT includes the SOR omega search)",
          file=sys.stderr, flush=True)
    warm = max(args.warmup, 2)
    d.algogas(sysp)
    tuning = d.autotune()            # launch shapes (the results do not depend on them)
    d.profile()                      # all kernels, to pick the dominant one
    for _ in range(warm - 1):
        d.algogas(sysp)
    warm_prof = d.profile_read()
    nit = tuning["sor_iters_per_launch"]
    kname = args.roofline_kernel
    if kname == "auto":
        # the kernel that moves the most algorithmic bytes per step (the
        # roofline is an HBM one; by time the three largest kernels are
        # within a few microseconds of each other)
        def bytes_per_step(k):
            per_launch = ALG_BYTES_PER_CELL.get(k, 0) * (nit if k == "k_sor_iter" else 1)
            return per_launch * warm_prof[k][1]
        kname = max(warm_prof, key=bytes_per_step)
    d.profile_off()

    def run(n):
        """n DTs between two device syncs: wall seconds, substeps, SOR
        iterations counted and computed"""
        sub, its, swept = 0, 0, 0
        d.sync()
        ta = time.perf_counter()
        for _ in range(n):
            lg = d.algogas(sysp)
            sub += lg.nsub
            its += sum(lg.sor_niter[k] for k in range(lg.nsub))
            # iterations the launches that swept computed (a solve that stops at
            # the first iteration of a two-iteration launch has computed the
            # second too)
            swept += sum((lg.sor_niter[k] + nit - 1) // nit * nit for k in range(lg.nsub))
        d.sync()
        return time.perf_counter() - ta, sub, its, swept

    # the timed region: the step as a caller runs it, no events between launches
    elapsed, substeps, sor_iters, sor_swept = run(args.steps)
    # the same number of DTs again with HIP events around the dominant
    # kernel's launches on the library's stream (events between launches
    # stretch the step by a few per cent, so the throughput above is taken
    # without them)
    d.profile(kernels=[kname])
    elapsed_ev, substeps_ev, sor_iters_ev, sor_swept_ev = run(args.steps)
    prof = d.profile_read()
    d.profile_off()
    value = args.steps / elapsed
    ms_k, n_k = prof.get(kname, (0.0, 0))
    avg_s = (ms_k / max(n_k, 1)) * 1e-3
    cells = (nrad - 2) * nsec if kname.startswith("k_sor") or kname == "k_matrix" else nrad * nsec
    alg = ALG_BYTES_PER_CELL.get(kname, 0) * cells
    # SOR iterations are queued ahead of the device-side stop test; the ones
    # queued after convergence return at once.  Only the iterations the
    # solves did move data: bytes = alg x iterations, time = all launches.
    # The first launch of each solve also assembles the energy system (the
    # matrix pass's compulsory bytes, once per substep).
    work = sor_iters_ev if kname == "k_sor_iter" else n_k
    form = ALG_BYTES_PER_CELL["k_matrix"] * cells * substeps_ev if kname == "k_sor_iter" else 0
    achieved = (alg * work + form) / (ms_k * 1e-3) / 1e9 if ms_k > 0 else 0.0
    swept = sor_swept_ev if kname == "k_sor_iter" else n_k
    achieved_swept = (alg * swept + form) / (ms_k * 1e-3) / 1e9 if ms_k > 0 else 0.0
    traffic = (measured_traffic(kname, nit if kname == "k_sor_iter" else None)
               if (nrad, nsec) == (1024, 1536) else None)
    step_bytes = (STEP_BYTES_PER_CELL * nrad * nsec +
                  SOR_BYTES_PER_CELL_ITER * (nrad - 2) * nsec * sor_iters / args.steps)
    out = {
        "metric": METRIC, "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": warm, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"damped_{nrad}x{nsec}", "nrad": nrad, "nsec": nsec, "dt": 1e-3,
                   "substeps": substeps, "sor_iters_per_step": sor_iters / args.steps,
                   "opacity": "LP85", "boundary": "DAMPING/ZEROVRAD", "planets": 1, "tuning": tuning},
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg * (nit if kname == "k_sor_iter" else 1),
                     "assembly_bytes_per_solve": (ALG_BYTES_PER_CELL["k_matrix"] * cells
                                                  if kname == "k_sor_iter" else 0),
                     "avg_launch_us": avg_s * 1e6, "launches": n_k,
                     "iterations_counted": work, "iterations_computed": swept,
                     "frac_iterations_computed": achieved_swept / HBM_PEAK_GBS,
                     "kernel_share_of_step": (ms_k * 1e-3) / elapsed_ev,
                     "timed_over": f"{args.steps} DTs after the throughput region, HIP events around each "
                                   f"launch of the kernel ({1e3 * elapsed_ev / args.steps:.4f} ms per DT there)"},
        "step_roofline": {"alg_bytes_per_step": step_bytes, "achieved": step_bytes / (elapsed / args.steps) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step_warmup": {k: v[0] / max(warm - 1, 1) for k, v in sorted(
            warm_prof.items(), key=lambda kv: -kv[1][0])},
    }
    d.close()
    print("[bench] the same workload stepped by the C driver", file=sys.stderr, flush=True)
    out["c_driver"] = c_driver_timing(nrad, nsec)
    if not args.no_sor_study:
        print("[bench] energy-solve study at 2048x3072", file=sys.stderr, flush=True)
        out["sor_study"] = sor_study()
    try:
        # in a process of its own: torch's HIP runtime and the library's do
        # not share one process here
        import subprocess
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-bandwidth"], capture_output=True,
                           text=True, timeout=300)
        out["hbm_copy"] = json.loads(r.stdout.strip().splitlines()[-1])
        out["roofline"]["frac_of_copy"] = out["roofline"]["achieved"] / out["hbm_copy"]["GBps"]
    except Exception as e:  # the figure is optional
        out["hbm_copy"] = {"error": str(e)}
    if not args.no_ensemble:
        print("[bench] ensemble of 16 disks", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_study()
    if not args.no_two_fluid:
        print("[bench] two-fluid workload at 1024x1536", file=sys.stderr, flush=True)
        out["two_fluid"] = two_fluid_study(nrad, nsec)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline_reference(nrad, nsec)
        except Exception as e:  # the GPU figure stands without it
            out["cpu_baseline"] = {"error": str(e)}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
