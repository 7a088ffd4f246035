"""CPU oracle of the FARGO_THORIN hydro step -- TEST INFRASTRUCTURE ONLY.

Python binding of oracle/liboracle.so (a sequential C restatement of the
reference's AlgoGas, oracle/thorin_oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from fargo_thorin_amd._capi import FIELD_ID, NBODY_FN, Planets, State, StepLog

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "thorin")
_lib = None

ONE_D = {"radii": 0, "rinf": 1, "rsup": 2, "rmed": 3, "surf": 4, "inv_diff_rmed": 5, "sigma_med": 6,
         "sigma_inf": 7, "energy_med": 8, "wave_killer": 9, "globcsvec": 10, "globtempervec": 11,
         "vtheta_med": 12, "nshift": 13, "vmed": 14, "peb_dens_init": 15, "peb_vrad_init": 16,
         "peb_vtheta_init": 17}


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, sz = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_size_t
    for name, res, args in [
        ("oracle_create", ip, [vp, C.POINTER(vp)]),
        ("oracle_destroy", None, [vp]),
        ("oracle_init_from_arrays", ip, [vp, dp, dp, dp, dp]),
        ("oracle_init_pebbles", ip, [vp, ip, dp, dp, dp, dp]),
        ("oracle_algogas", ip, [vp, C.POINTER(Planets), NBODY_FN, vp, C.POINTER(StepLog)]),
        ("oracle_get_field", ip, [vp, ip, dp, sz]),
        ("oracle_set_field", ip, [vp, ip, dp, sz]),
        ("oracle_get_state", ip, [vp, C.POINTER(State)]),
        ("oracle_set_state", ip, [vp, C.POINTER(State)]),
        ("oracle_get_1d", ip, [vp, ip, dp, sz]),
        ("oracle_gas_total_mass", C.c_double, [vp]),
        ("oracle_condition_cfl", ip, [vp, C.c_double, C.c_double, dp]),
        ("oracle_opacity_law", C.c_double, [ip, C.c_double, C.c_double, C.c_double]),
        ("oracle_flux_limiter", C.c_double, [C.c_double]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class OracleDisk:
    """Same shape as fargo_thorin_amd.Disk, computed by the CPU oracle."""

    def __init__(self, params):
        h = C.c_void_p()
        if lib().oracle_create(params.ptr, C.byref(h)) != 0:
            raise RuntimeError("oracle_create failed")
        self._h = h
        self.nrad = int(params.get("NRAD"))
        self.nsec = int(params.get("NSEC"))
        self.log = StepLog()

    def init(self, dens, vrad, vtheta, temper):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (dens, vrad, vtheta, temper)]
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        if lib().oracle_init_from_arrays(self._h, *ptrs) != 0:
            raise RuntimeError("oracle init failed")

    def init_pebbles(self, mode: int = 0, pdens=None, pvrad=None, pvtheta=None, psize=None):
        """mode 0: equilibrium pebble disk from the gas; mode 1: restart from arrays."""
        if mode == 0:
            ptrs = [None] * 4
        else:
            arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (pdens, pvrad, pvtheta, psize)]
            self._peb_keep = arrs
            ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        if lib().oracle_init_pebbles(self._h, mode, *ptrs) != 0:
            raise RuntimeError("oracle pebble init failed")

    def algogas(self, planets: Planets, nbody) -> StepLog:
        rc = lib().oracle_algogas(self._h, C.byref(planets), nbody, None, C.byref(self.log))
        if rc < 0:
            raise RuntimeError(f"oracle_algogas failed ({rc})")
        return self.log

    def field(self, name: str, ghost: bool = False) -> np.ndarray:
        rings = self.nrad + (1 if ghost else 0)
        out = np.empty((rings, self.nsec))
        fid = FIELD_ID[name] if isinstance(name, str) else int(name)
        lib().oracle_get_field(self._h, fid, out.ctypes.data_as(C.POINTER(C.c_double)), out.size)
        return out

    def set_field(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        lib().oracle_set_field(self._h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)), a.size)

    def one_d(self, which: str) -> np.ndarray:
        out = np.empty(self.nrad + 1)
        lib().oracle_get_1d(self._h, ONE_D[which], out.ctypes.data_as(C.POINTER(C.c_double)), out.size)
        return out

    def state(self) -> State:
        s = State()
        lib().oracle_get_state(self._h, C.byref(s))
        return s

    def set_state(self, s: State) -> None:
        lib().oracle_set_state(self._h, C.byref(s))

    def total_mass(self) -> float:
        return lib().oracle_gas_total_mass(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
