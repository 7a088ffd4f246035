"""Host side of the library without a GPU: the C ABI exports every entry
point include/thorin_amd.h declares, the in.par reader keeps the reference's
defaults and derived flags, out-of-scope options are refused, and the
host-only set-up (geometry, initial state, pebble disk) equals the oracle and
the reference's own output bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fargo_thorin_amd import Params, _capi, configs, seeded_disk, seeded_pebbles
from fargo_thorin_amd._capi import Geometry, lib, last_error
from oracle import OracleDisk
from tests.helpers import case, golden, params_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thorin_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(thorin_[a-z0-9_]+)\s*\(", text))
    # typedef'd callback types are not entry points
    return sorted(n for n in names if n != "thorin_nbody_fn")


def test_library_exports_every_declared_entry_point():
    L = C.CDLL(_capi.LIB_PATH if hasattr(_capi, "LIB_PATH") else lib()._name)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_params_defaults_and_derived_flags():
    P = Params(ThicknessSmoothing="0.6")   # var.c leaves it 0, which Interpret.c refuses
    # var.c defaults apply to optional keys only (Interpret.c:52-60): the
    # necessary ones (DT, NRAD, ...) stay 0 until the parameter file sets them
    assert P.num("DT") == 0.0 and P.num("NRAD") == 0
    assert P.num("ADIABIND") == 1.4 and P.num("PEBBLEALPHA") == 1e-4 and P.num("EVAPORATIONRATE") == 1e-3
    # Interpret.c:241 assigns instead of comparing: particle diffusion is always on
    assert P.flag("diffusive_particles") == 1
    Q = Params(ThicknessSmoothing="0.6", PebbleAccretion="yes", BackReaction="NO", Transport="STANDARD")
    assert Q.flag("pebbles") == 1 and Q.flag("back_reaction") == 0 and Q.flag("fast_transport") == 0


@pytest.mark.parametrize("keys,msg", [
    (dict(InnerBoundary="OPEN"), "OPEN"),
    (dict(Frame="COROTATING"), "COROTATING"),
    (dict(EnergyEquation="NO"), "isothermal"),
    (dict(Nsec="63"), "NSEC"),
])
def test_out_of_scope_options_are_refused(keys, msg):
    P = params_for(configs.with_grid(configs.PLAIN_128x256, 16, 32), **keys)
    h = C.c_void_p()
    rc = lib().thorin_ctx_create_host(P.ptr, C.byref(h))
    assert rc != 0 and msg.lower() in last_error().lower()


def test_sm03_opacity_needs_its_table(tmp_path):
    """OpacityTable SM03 reads INPUTDIR/data_gasopacity_malygin.txt at set-up
    (InitOpacities, Opacity2.c:252-300): a missing file is an error that names
    it; with the table the context is made."""
    cfg = configs.with_grid(configs.DAMPED_1024x1536, 16, 32, OpacityTable="SM03")
    h = C.c_void_p()
    P = params_for(cfg, InputDir=str(tmp_path) + "/")
    assert lib().thorin_ctx_create_host(P.ptr, C.byref(h)) != 0
    assert "data_gasopacity_malygin.txt" in last_error() and "not found" in last_error()
    (tmp_path / "data_gasopacity_malygin.txt").write_text("126\n94\n1.0\n")
    assert lib().thorin_ctx_create_host(P.ptr, C.byref(h)) != 0 and "malformed" in last_error()
    P = params_for(cfg)                    # the fixture table
    h = host_ctx(P)
    assert P.get("OPACITYTABLE")[:1] == "S"
    lib().thorin_ctx_destroy(h)


def host_ctx(P):
    h = C.c_void_p()
    assert lib().thorin_ctx_create_host(P.ptr, C.byref(h)) == 0, last_error()
    return h


def host_field(h, fid, nr, ns):
    out = np.empty((nr, ns))
    assert lib().thorin_host_field(h, fid, out.ctypes.data_as(C.POINTER(C.c_double)), out.size) == 0, last_error()
    return out


def test_host_geometry_and_initial_state_equal_the_oracle():
    cfg = configs.with_grid(configs.DAMPED_1024x1536, 40, 64)
    P = params_for(cfg)
    disk = seeded_disk(40, 64, seed=5)
    h = host_ctx(P)
    try:
        arrs = [np.ascontiguousarray(disk[k]).ravel() for k in ("dens", "vrad", "vtheta", "temper")]
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        assert lib().thorin_init_state(h, 1, *ptrs) == 0
        geo = Geometry()
        lib().thorin_get_geometry(h, C.byref(geo))
        o = OracleDisk(P)
        o.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        for name, key in (("radii", "radii"), ("rinf", "rinf"), ("rmed", "rmed"), ("surf", "surf"),
                          ("inv_diff_rmed", "inv_diff_rmed"), ("sigma_med", "sigma_med"),
                          ("wave_killer", "wave_killer")):
            a = np.ctypeslib.as_array(getattr(geo, name), shape=(40,))
            assert np.array_equal(a, o.one_d(key)[:40]), name
        for f in ("dens", "energy", "soundspeed", "temper", "pressure"):
            assert np.array_equal(host_field(h, _capi.FIELD_ID[f], 40, 64), o.field(f)), f
    finally:
        lib().thorin_ctx_destroy(h)


def test_host_pebble_equilibrium_equals_reference_output():
    """EquilPebbleDisk of the library's host code against the reference's
    gaspebble*0.dat of the two-fluid golden run, bit for bit."""
    name = "pebbles_q3e-5_64x128"
    cfg, _, _ = case(name)
    g = golden(name)
    h = host_ctx(params_for(cfg))
    try:
        arrs = [np.ascontiguousarray(g[k]).ravel() for k in ("init_dens", "init_vrad", "init_vtheta", "init_temper")]
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        assert lib().thorin_init_state(h, 1, *ptrs) == 0
        assert lib().thorin_init_pebbles(h, 0, None, None, None, None) == 0, last_error()
        for f in ("pebbledens", "pebblevrad", "pebblevtheta", "pebblesize", "pebblestokes"):
            assert np.array_equal(host_field(h, _capi.FIELD_ID[f], 64, 128), g[f"{f}0"]), f
    finally:
        lib().thorin_ctx_destroy(h)


def test_host_pebble_restart_targets_equal_the_oracle():
    """RestartPebbleDisk + BckpFieldsForBC: the damping targets are the
    ordered ring means (restated independently by the oracle)."""
    cfg = configs.with_grid(configs.PEBBLES_1024x1536, 32, 48)
    P = params_for(cfg)
    disk = seeded_disk(32, 48, seed=9)
    pb = seeded_pebbles(disk)
    o = OracleDisk(P)
    o.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    o.init_pebbles(1, pb["pdens"], pb["pvrad"], pb["pvtheta"], pb["psize"])
    # the means of ring-constant rows are the row value up to the rounding of
    # the ordered sum: check against numpy's sequential sum
    for key, arr, add in (("peb_dens_init", pb["pdens"], 0.0), ("peb_vrad_init", pb["pvrad"], 0.0),
                          ("peb_vtheta_init", pb["pvtheta"], 0.0)):
        seq = np.array([sum(row.tolist()) for row in arr])
        assert np.array_equal(o.one_d(key)[:32], seq / 48.0 + add), key
    h = host_ctx(P)
    try:
        arrs = [np.ascontiguousarray(disk[k]).ravel() for k in ("dens", "vrad", "vtheta", "temper")]
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        assert lib().thorin_init_state(h, 1, *ptrs) == 0
        parr = [np.ascontiguousarray(pb[k]).ravel() for k in ("pdens", "pvrad", "pvtheta", "psize")]
        pp = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in parr]
        assert lib().thorin_init_pebbles(h, 1, *pp) == 0, last_error()
        assert np.array_equal(host_field(h, _capi.FIELD_ID["pebbledens"], 32, 48), pb["pdens"])
    finally:
        lib().thorin_ctx_destroy(h)


def test_ensemble_refuses_contexts_without_a_device():
    """thorin_ensemble_create checks its contexts before touching the GPU."""
    P = params_for(configs.with_grid(configs.PLAIN_128x256, 16, 32))
    h = host_ctx(P)
    try:
        e = C.c_void_p()
        arr = (C.c_void_p * 1)(h)
        assert lib().thorin_ensemble_create(arr, 1, C.byref(e)) == -1
        assert "no device" in last_error()
        assert lib().thorin_ensemble_create(arr, 0, C.byref(e)) == -1
        assert lib().thorin_ensemble_create(None, 1, C.byref(e)) == -1
    finally:
        lib().thorin_ctx_destroy(h)


def test_ensemble_members_follow_the_workload_draws():
    """16 members: Nrad from the five sizes, Nsec = 1.5 Nrad, alpha in
    [1e-4, 1e-2] log-uniform, Sigma0 in 5e-4 x [0.3, 3], h in [0.03, 0.07];
    seeds 1000+d; every member is a valid parameter set."""
    cfgs = configs.ensemble_configs()
    assert len(cfgs) == 16 and [c["_seed"] for c in cfgs] == list(range(1000, 1016))
    for c in cfgs:
        nr, ns = int(c["Nrad"]), int(c["Nsec"])
        assert nr in configs.ENSEMBLE_SIZES and ns * 2 == 3 * nr
        assert 1e-4 <= float(c["AlphaViscosity"]) <= 1e-2
        assert 0.3 * 5e-4 <= float(c["Sigma0"]) <= 3.0 * 5e-4
        assert 0.03 <= float(c["AspectRatio"]) <= 0.07
        P = params_for({k: v for k, v in c.items() if k != "_seed"})
        assert P.num("NRAD") == nr
    # the draws differ between members
    assert len({c["AlphaViscosity"] for c in cfgs}) == 16
