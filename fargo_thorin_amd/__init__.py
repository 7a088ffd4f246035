"""fargo_thorin_amd -- MI355X-native hydro step (AlgoGas) of FARGO_THORIN.

The compute path is libthorin_amd.so (HIP kernels for gfx950 behind the C ABI
of include/thorin_amd.h).  This module is a thin ctypes binding that mirrors
the reference's host interface:

    Params      in.par parameters (ReadVariables, src_main/Interpret.c:65-266)
    Disk        the PolarGrid fields resident on the GPU; Disk.algogas() is
                one call of AlgoGas (src_main/SourceEuler.c:210-317)
    Planets     the planetary system (types.h:96-117)

Nothing here computes physics on the CPU; without the built library every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _capi
from ._capi import FIELD_ID, FIELDS, OPACITY, Planets, State, StepLog, last_error, lib
from .seed import seeded_disk, seeded_pebbles, write_cfg

__all__ = ["Params", "Disk", "Planets", "State", "StepLog", "seeded_disk", "write_cfg", "FIELDS",
           "OPACITY", "ThorinError", "make_planets", "WarmStart", "seeded_pebbles", "Ensemble"]


class ThorinError(RuntimeError):
    pass


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise ThorinError(f"{what} failed: {last_error()}")
    return rc


class Params:
    """in.par parameters with the reference's defaults (var.c:19-139)."""

    def __init__(self, path: Optional[str] = None, **keys):
        L = lib()
        self._p = L.thorin_params_new()
        if not self._p:
            raise ThorinError("out of memory")
        if path is not None:
            _check(L.thorin_params_read(path.encode(), self._p), f"reading {path}")
        for k, v in keys.items():
            self.set(k, v)
        self.derive()

    def set(self, key: str, value) -> "Params":
        _check(lib().thorin_params_set(self._p, key.encode(), str(value).encode()), f"set {key}")
        return self

    def derive(self) -> "Params":
        _check(lib().thorin_params_derive(self._p), "parameter check")
        return self

    def get(self, key: str) -> str:
        buf = C.create_string_buffer(1024)
        _check(lib().thorin_params_get(self._p, key.encode(), buf, 1024), f"get {key}")
        return buf.value.decode()

    def num(self, key: str) -> float:
        return float(self.get(key))

    def flag(self, name: str) -> int:
        return lib().thorin_params_flag(self._p, name.encode())

    @property
    def ptr(self):
        return self._p

    def __del__(self):
        try:
            if getattr(self, "_p", None):
                lib().thorin_params_free(self._p)
                self._p = None
        except Exception:
            pass


def make_planets(masses, radii_au, feel_disk: bool = False) -> Planets:
    """Circular, coplanar planets at phase 0 (the state REBOUND's
    reb_tools_orbit_to_particle gives for e = inc = f = 0)."""
    s = Planets()
    s.nb = len(masses)
    for k, (m, a) in enumerate(zip(masses, radii_au)):
        m = float(np.float32(m))   # the planet file is parsed as float (ReboundInterface.c:162)
        a = float(np.float32(a))
        s.mass[k] = m
        s.x[k] = a
        s.vy[k] = (1.0 * (1.0 + m) / a) ** 0.5
        s.feeldisk[k] = int(feel_disk)
    return s


def _tuning_dict(t) -> dict:
    return {"sor_iters_per_launch": t.sor_iters_per_launch, "rad_rows": t.rad_rows,
            "sor_us_per_iter": list(t.sor_us_per_iter), "rad_us": list(t.rad_us), "fused_chain": t.fused_chain,
            "sor_assembly": t.sor_assembly}


class Disk:
    """All PolarGrid fields of one disk, resident on the GPU."""

    def __init__(self, params: Params):
        self.params = params
        L = lib()
        h = C.c_void_p()
        _check(L.thorin_ctx_create(params.ptr, C.byref(h)), "thorin_ctx_create")
        self._h = h
        self.nrad = int(params.get("NRAD"))
        self.nsec = int(params.get("NSEC"))
        self.log = StepLog()
        self._nbody = _capi.host_nbody_fn()

    # --- initialisation (Init.c:72-145)
    def init(self, dens=None, vrad=None, vtheta=None, temper=None, mode: Optional[int] = None) -> None:
        if dens is None:
            _check(lib().thorin_init_state(self._h, 0, None, None, None, None), "init")
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (dens, vrad, vtheta, temper)]
        n = self.nrad * self.nsec
        for a in arrs:
            if a.size != n:
                raise ValueError(f"initial arrays must hold NRAD*NSEC = {n} values")
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        _check(lib().thorin_init_state(self._h, 1 if mode is None else mode, *ptrs), "init")

    def init_pebbles(self, mode: int = 0, pdens=None, pvrad=None, pvtheta=None, psize=None) -> None:
        """Pebble disk (Init.c:138-144): mode 0 = coagulation-drift equilibrium
        from the initial gas (EquilPebbleDisk); mode 1 = the given pebble
        density, velocities and sizes (RestartPebbleDisk)."""
        if mode == 0:
            _check(lib().thorin_init_pebbles(self._h, 0, None, None, None, None), "init_pebbles")
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (pdens, pvrad, pvtheta, psize)]
        n = self.nrad * self.nsec
        for a in arrs:
            if a.size != n:
                raise ValueError(f"pebble arrays must hold NRAD*NSEC = {n} values")
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        _check(lib().thorin_init_pebbles(self._h, mode, *ptrs), "init_pebbles")

    # --- the hot path
    def algogas(self, planets: Planets, nbody=None) -> StepLog:
        """One DT of AlgoGas; returns the step log (substeps, dt, SOR iterations)."""
        cb = self._nbody if nbody is None else nbody
        rc = lib().thorin_algogas(self._h, C.byref(planets), cb, None, C.byref(self.log))
        _check(rc, "thorin_algogas")
        self.last_rc = rc
        return self.log

    # --- launch shapes (results do not depend on them)
    def autotune(self) -> dict:
        """Time the launch shapes on this grid and keep the fastest."""
        t = _capi.Tuning()
        _check(lib().thorin_autotune(self._h, C.byref(t)), "thorin_autotune")
        return _tuning_dict(t)

    def set_tuning(self, sor_iters_per_launch: int = 3, rad_rows: int = 8, fused_chain: int = 1,
                   sor_assembly: int = 1) -> None:
        t = _capi.Tuning(sor_iters_per_launch=sor_iters_per_launch, rad_rows=rad_rows, fused_chain=fused_chain,
                         sor_assembly=sor_assembly)
        _check(lib().thorin_set_tuning(self._h, C.byref(t)), "thorin_set_tuning")

    def tuning(self) -> dict:
        t = _capi.Tuning()
        _check(lib().thorin_get_tuning(self._h, C.byref(t)), "thorin_get_tuning")
        return _tuning_dict(t)

    # --- fields
    def field(self, name: str, ghost: bool = False) -> np.ndarray:
        rings = self.nrad + (1 if ghost else 0)
        out = np.empty((rings, self.nsec), dtype=np.float64)
        _check(lib().thorin_download_field(self._h, FIELD_ID[name], out.ctypes.data_as(C.POINTER(C.c_double)),
                                           out.size), f"download {name}")
        return out

    def upload(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(lib().thorin_upload_field(self._h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)),
                                         a.size), f"upload {name}")

    def ring_means(self, name: str) -> np.ndarray:
        out = np.empty(self.nrad)
        _check(lib().thorin_ring_means(self._h, FIELD_ID[name], out.ctypes.data_as(C.POINTER(C.c_double))),
               "ring means")
        return out

    def ring_profile(self, which: str) -> np.ndarray:
        idx = {"globcsvec": 0, "globtempervec": 1, "vmoy": 2, "vmed": 3, "nshift": 4}[which]
        out = np.empty(self.nrad)
        _check(lib().thorin_ring_profile(self._h, idx, out.ctypes.data_as(C.POINTER(C.c_double))), "ring profile")
        return out

    def total_mass(self) -> float:
        v = C.c_double()
        _check(lib().thorin_gas_total_mass(self._h, C.byref(v)), "total mass")
        return v.value

    def state(self) -> State:
        s = State()
        _check(lib().thorin_get_state(self._h, C.byref(s)), "state")
        return s

    def set_state(self, s: State) -> None:
        _check(lib().thorin_set_state(self._h, C.byref(s)), "state")

    def geometry(self) -> dict:
        g = _capi.Geometry()
        _check(lib().thorin_get_geometry(self._h, C.byref(g)), "geometry")
        n = self.nrad + 1
        out = {}
        for name, _ in _capi.Geometry._fields_[2:]:
            p = getattr(g, name)
            out[name] = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return out

    # --- kernel timing (HIP events on the library's stream)
    @staticmethod
    def kernel_names() -> list:
        out, k = [], 0
        while True:
            n = lib().thorin_prof_kernel_name(k)
            if not n:
                return out
            out.append(n.decode())
            k += 1

    def profile(self, kernels=None) -> None:
        names = self.kernel_names()
        mask = 0
        for k, n in enumerate(names):
            if kernels is None or n in kernels:
                mask |= 1 << k
        _check(lib().thorin_prof_enable(self._h, 1, mask), "profile")

    def profile_off(self) -> None:
        _check(lib().thorin_prof_enable(self._h, 0, 0), "profile")

    def profile_read(self) -> dict:
        out = {}
        for k, n in enumerate(self.kernel_names()):
            ms, cnt = C.c_double(), C.c_longlong()
            _check(lib().thorin_prof_read(self._h, k, C.byref(ms), C.byref(cnt)), "profile read")
            if cnt.value:
                out[n] = (ms.value, cnt.value)
        return out

    # --- the energy equation's implicit solve on its own
    def rad_system(self, dt: float = 0.0) -> float:
        """Assemble SubStep3's radiative-diffusion system of the current fields
        for a step dt (dt <= 0: the hydro CFL step); its T is kept as the
        initial guess of rad_solve.  Returns the dt used."""
        used = C.c_double()
        _check(lib().thorin_rad_system(self._h, C.c_double(dt), C.byref(used)), "thorin_rad_system")
        return used.value

    def rad_solve(self, method: int = 0, tol: float = 1e-10, omega: float = 1.8, inner_sweeps: int = 8,
                  max_iters: int = 100000, warm: "Optional[WarmStart]" = None, sweep_block: int = 0,
                  noalloc: bool = False) -> dict:
        """Solve the assembled system from its initial guess: method 0 = FP64
        red-black SOR, 1 = FP32 SOR sweeps in FP64 iterative refinement; with
        warm (method 1), from the network's guess when it lowers the residual.
        noalloc: the allocation-free entry point (thorin_rad_solve_noalloc).
        T is left in field "temper"."""
        o = _capi.SolveOpts(method=method, inner_sweeps=inner_sweeps, max_iters=max_iters, tol=tol, omega=omega,
                            warm=warm._h if warm is not None else None, sweep_block=sweep_block)
        r = _capi.SolveResult()
        fn = lib().thorin_rad_solve_noalloc if noalloc else lib().thorin_rad_solve
        _check(fn(self._h, C.byref(o), C.byref(r)), "thorin_rad_solve_noalloc" if noalloc else "thorin_rad_solve")
        return {"sweeps": r.sweeps, "rounds": r.rounds, "converged": bool(r.converged), "ms": r.ms,
                "resid0": r.resid0, "resid": r.resid, "warm_used": bool(r.warm_used),
                "resid_warm": r.resid_warm}

    def sync(self) -> None:
        _check(lib().thorin_sync(self._h), "sync")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().thorin_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Disks advanced together (thorin_ensemble_*): every kernel of a substep
    is one launch over all disks' grids; each disk's results are those of its
    own Disk.algogas, bit for bit."""

    def __init__(self, disks):
        self.disks = list(disks)
        n = len(self.disks)
        arr = (C.c_void_p * n)(*[d._h for d in self.disks])
        h = C.c_void_p()
        _check(lib().thorin_ensemble_create(arr, n, C.byref(h)), "thorin_ensemble_create")
        self._h = h
        self.logs = (StepLog * n)()
        self.rcs = (C.c_int * n)()
        self._nbody = _capi.host_nbody_fn()

    def algogas(self, planets, nbody=None):
        """One DT of every disk; planets: one Planets per disk.  Returns the
        per-disk step logs."""
        n = len(self.disks)
        if len(planets) != n:
            raise ValueError(f"one Planets per disk is needed ({n})")
        ptrs = (C.POINTER(Planets) * n)(*[C.pointer(p) for p in planets])
        cb = self._nbody if nbody is None else nbody
        rc = lib().thorin_ensemble_algogas(self._h, ptrs, cb, None, self.logs, self.rcs)
        _check(rc, "thorin_ensemble_algogas")
        self.last_rc = rc
        return list(self.logs)

    def stats(self):
        """(launches issued, per-disk kernel operations they carried, graphs
        captured, graph replays)"""
        v = [C.c_longlong() for _ in range(4)]
        _check(lib().thorin_ensemble_stats(self._h, *[C.byref(x) for x in v]), "thorin_ensemble_stats")
        return tuple(x.value for x in v)

    def autotune(self) -> dict:
        """One launch shape for all disks: the least time summed over them."""
        t = _capi.Tuning()
        _check(lib().thorin_ensemble_autotune(self._h, C.byref(t)), "thorin_ensemble_autotune")
        return _tuning_dict(t)

    def set_graphs(self, on: bool) -> None:
        """Replay each repeating merged substep as a HIP graph (default on)."""
        _check(lib().thorin_ensemble_set_graphs(self._h, int(bool(on))), "thorin_ensemble_set_graphs")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().thorin_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmStart:
    """The learned warm start of the energy solve (thorin_cnn_*): a 4-level,
    32-channel U-Net trained on the GPU on pairs the solver generates."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        _check(lib().thorin_cnn_create(seed, C.byref(h)), "thorin_cnn_create")
        self._h = h

    def n_params(self) -> int:
        return int(lib().thorin_cnn_param_count(self._h))

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params(), dtype=np.float32)
        _check(lib().thorin_cnn_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size), "params")
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().thorin_cnn_set_params(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "set params")

    def grad(self, x, t=None, aux=None):
        """forward + backward on host arrays: (loss, gradient, prediction);
        with aux (5 x H x W, the system's grid) the residual loss.  x, t and
        the prediction are on the network's grid, H/f x W/f."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        f = self.coarsen()
        Hn, Wn = x.shape[:2]
        H, W = Hn * f, Wn * f
        fp = C.POINTER(C.c_float)
        tp = ap = None
        if t is not None:
            t = np.ascontiguousarray(t, dtype=np.float32)
            tp = t.ctypes.data_as(fp)
        if aux is not None:
            aux = np.ascontiguousarray(aux, dtype=np.float32)
            ap = aux.ctypes.data_as(fp)
        g = np.empty(self.n_params(), dtype=np.float32)
        y = np.empty((Hn, Wn), dtype=np.float32)
        L = C.c_double()
        _check(lib().thorin_cnn_grad2(self._h, H, W, x.ctypes.data_as(fp), tp, ap, C.byref(L),
                                      g.ctypes.data_as(fp), y.ctypes.data_as(fp)), "thorin_cnn_grad")
        return L.value, g, y

    def set_loss(self, mode: int) -> None:
        """0: error of the normalised correction; 1: residual of the guess"""
        _check(lib().thorin_cnn_set_loss(self._h, mode), "set_loss")

    def set_coarsen(self, f: int) -> None:
        """run the network on the system's grid coarsened f = 1, 2 or 4 times
        (before the first pair)"""
        _check(lib().thorin_cnn_set_coarsen(self._h, f), "set_coarsen")

    def coarsen(self) -> int:
        return int(lib().thorin_cnn_coarsen(self._h))

    def add_pair(self, disk: Disk) -> None:
        _check(lib().thorin_cnn_add_pair(self._h, disk._h), "thorin_cnn_add_pair")

    def n_pairs(self) -> int:
        return int(lib().thorin_cnn_num_pairs(self._h))

    def train(self, iters: int, lr: float = 1e-3):
        a, b = C.c_double(), C.c_double()
        _check(lib().thorin_cnn_train(self._h, iters, lr, C.byref(a), C.byref(b)), "thorin_cnn_train")
        return a.value, b.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().thorin_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
