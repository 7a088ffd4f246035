"""ctypes declarations of the C ABI in include/thorin_amd.h.

The structures here mirror the header field for field; tests check the sizes
against the library (thorin_struct_sizes).
"""
from __future__ import annotations

import ctypes as C
import os

MAXPLANETS = 120
MAXSUB = 4096
STRLEN = 512

# field ids (thorin_amd.h)
FIELDS = ["dens", "vrad", "vtheta", "energy", "label", "temper", "pressure", "soundspeed", "qplus",
          "qminus", "opacity", "divv", "taurr", "taurp", "taupp", "agr", "agt", "energyint", "energynew",
          "vradint", "vthetaint", "vradnew", "vthetanew", "qbalance", "pebbledens", "pebblevrad",
          "pebblevtheta", "pebblesize", "pebblestokes", "eta", "torque"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
# read-only diagnostics: the energy equation's five-point system (THORIN_X_*)
FIELD_ID.update({"sys_ml": 64, "sys_mlip": 65, "sys_mlim": 66, "sys_mljp": 67, "sys_mljm": 68, "sys_rhs": 69,
                 "sys_diff": 70,
                 # the first crash's fields (the gas*999.dat dumps)
                 "crash_dens": 71, "crash_vrad": 72, "crash_vtheta": 73, "crash_temper": 74})
OPACITY = {"BL94": 0, "LP85": 1, "ZHU12": 2, "CONST": 3, "POWERLAW": 4}


class Planets(C.Structure):
    _fields_ = [("nb", C.c_int)] + [(n, C.c_double * MAXPLANETS) for n in
                                      ("mass", "gas", "solid", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "acc")] + \
               [("feeldisk", C.c_int * MAXPLANETS)]


class State(C.Structure):
    _fields_ = [("physical_time", C.c_double), ("physical_time_initial", C.c_double),
                ("omega_frame", C.c_double), ("mass_taper", C.c_double), ("dt_stellar", C.c_double),
                ("sor_omega", C.c_double), ("sor_domega", C.c_double), ("sor_first", C.c_int),
                ("sor_niter_last", C.c_int), ("sor_niter_best", C.c_int), ("sor_omega_best", C.c_double),
                ("cfl_count", C.c_int), ("already_crashed", C.c_int)]


class StepLog(C.Structure):
    _fields_ = [("nsub", C.c_int), ("crashed_any", C.c_int), ("cfl_n", C.c_int * MAXSUB),
                ("dt", C.c_double * MAXSUB), ("sor_niter", C.c_int * MAXSUB),
                ("sor_omega", C.c_double * MAXSUB), ("crashed", C.c_int * MAXSUB), ("lost_mass", C.c_double),
                ("crash_dumped", C.c_int), ("crash_time", C.c_double),
                ("tq_rows", C.c_int), ("tq", (C.c_double * 7) * 8)]


class Geometry(C.Structure):
    _fields_ = [("nrad", C.c_int), ("nsec", C.c_int)] + [(n, C.POINTER(C.c_double)) for n in (
        "radii", "rinf", "rsup", "rmed", "surf", "inv_rmed", "inv_surf", "inv_diff_rmed", "inv_diff_rsup",
        "inv_rinf", "omega_inv", "rmed2", "sigma_med", "sigma_inf", "energy_med", "vtheta_med", "wave_killer",
        "globcsvec", "globtempervec")]


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("inner_sweeps", C.c_int), ("max_iters", C.c_int),
                ("tol", C.c_double), ("omega", C.c_double), ("warm", C.c_void_p), ("sweep_block", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("rounds", C.c_int), ("converged", C.c_int),
                ("ms", C.c_double), ("resid0", C.c_double), ("resid", C.c_double),
                ("warm_used", C.c_int), ("resid_warm", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("sor_iters_per_launch", C.c_int), ("rad_rows", C.c_int),
                ("sor_us_per_iter", C.c_double * 3), ("rad_us", C.c_double * 3), ("fused_chain", C.c_int),
                ("sor_assembly", C.c_int)]


NBODY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Planets), C.c_double, C.c_double, C.c_double)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libthorin_amd.so")
_lib = None


def lib() -> C.CDLL:
    """Loads the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C fargo_thorin_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, sz = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_size_t
    sig = {
        "thorin_params_new": (vp, []),
        "thorin_params_free": (None, [vp]),
        "thorin_params_set": (ip, [vp, C.c_char_p, C.c_char_p]),
        "thorin_params_get": (ip, [vp, C.c_char_p, C.c_char_p, ip]),
        "thorin_params_flag": (ip, [vp, C.c_char_p]),
        "thorin_params_derive": (ip, [vp]),
        "thorin_params_read": (ip, [C.c_char_p, vp]),
        "thorin_params_defaults": (ip, [vp]),
        "thorin_ctx_create": (ip, [vp, C.POINTER(vp)]),
        "thorin_ctx_destroy": (ip, [vp]),
        "thorin_get_geometry": (ip, [vp, C.POINTER(Geometry)]),
        "thorin_get_state": (ip, [vp, C.POINTER(State)]),
        "thorin_set_state": (ip, [vp, C.POINTER(State)]),
        "thorin_init_state": (ip, [vp, ip, dp, dp, dp, dp]),
        "thorin_init_pebbles": (ip, [vp, ip, dp, dp, dp, dp]),
        "thorin_upload_field": (ip, [vp, ip, dp, sz]),
        "thorin_download_field": (ip, [vp, ip, dp, sz]),
        "thorin_algogas": (ip, [vp, C.POINTER(Planets), NBODY_FN, vp, C.POINTER(StepLog)]),
        "thorin_ring_means": (ip, [vp, ip, dp]),
        "thorin_ensemble_create": (ip, [C.POINTER(vp), ip, C.POINTER(vp)]),
        "thorin_ensemble_algogas": (ip, [vp, C.POINTER(C.POINTER(Planets)), NBODY_FN, C.POINTER(vp),
                                         C.POINTER(StepLog), C.POINTER(ip)]),
        "thorin_ensemble_stats": (ip, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                       C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "thorin_ensemble_set_graphs": (ip, [vp, ip]),
        "thorin_autotune": (ip, [vp, C.POINTER(Tuning)]),
        "thorin_set_tuning": (ip, [vp, C.POINTER(Tuning)]),
        "thorin_get_tuning": (ip, [vp, C.POINTER(Tuning)]),
        "thorin_ensemble_autotune": (ip, [vp, C.POINTER(Tuning)]),
        "thorin_ensemble_destroy": (None, [vp]),
        "thorin_ring_profile": (ip, [vp, ip, dp]),
        "thorin_ctx_create_host": (ip, [vp, C.POINTER(vp)]),
        "thorin_host_field": (ip, [vp, ip, dp, sz]),
        "thorin_gas_total_mass": (ip, [vp, dp]),
        "thorin_sync": (ip, [vp]),
        "thorin_prof_enable": (ip, [vp, ip, C.c_ulonglong]),
        "thorin_prof_read": (ip, [vp, ip, dp, C.POINTER(C.c_longlong)]),
        "thorin_prof_kernel_name": (C.c_char_p, [ip]),
        "thorin_last_error": (C.c_char_p, []),
        "thorin_planets_from_config": (ip, [vp, C.POINTER(Planets)]),
        "thorin_nbody_step": (ip, [vp, C.POINTER(Planets), C.c_double, C.c_double, C.c_double]),
        "thorin_rad_system": (ip, [vp, C.c_double, dp]),
        "thorin_rad_solve": (ip, [vp, C.POINTER(SolveOpts), C.POINTER(SolveResult)]),
        "thorin_rad_solve_noalloc": (ip, [vp, C.POINTER(SolveOpts), C.POINTER(SolveResult)]),
        "thorin_math_eval": (ip, [ip, dp, dp, dp, sz]),
        "thorin_opacity_eval": (ip, [vp, dp, dp, dp, sz]),
        "thorin_orbit_elements": (ip, [C.POINTER(Planets), ip, C.c_double, dp]),
        "thorin_cnn_create": (ip, [C.c_ulonglong, C.POINTER(vp)]),
        "thorin_cnn_destroy": (ip, [vp]),
        "thorin_cnn_param_count": (sz, [vp]),
        "thorin_cnn_get_params": (ip, [vp, C.POINTER(C.c_float), sz]),
        "thorin_cnn_set_params": (ip, [vp, C.POINTER(C.c_float), sz]),
        "thorin_cnn_add_pair": (ip, [vp, vp]),
        "thorin_cnn_num_pairs": (ip, [vp]),
        "thorin_cnn_train": (ip, [vp, ip, C.c_float, dp, dp]),
        "thorin_cnn_grad": (ip, [vp, ip, ip, C.POINTER(C.c_float), C.POINTER(C.c_float), dp,
                                 C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "thorin_cnn_grad2": (ip, [vp, ip, ip, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), dp,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "thorin_cnn_set_loss": (ip, [vp, ip]),
        "thorin_cnn_set_coarsen": (ip, [vp, ip]),
        "thorin_cnn_coarsen": (ip, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return lib().thorin_last_error().decode(errors="replace")


def host_nbody_fn() -> NBODY_FN:
    """The library's own host N-body step as a callback object."""
    addr = C.cast(lib().thorin_nbody_step, C.c_void_p).value
    return NBODY_FN(addr)
