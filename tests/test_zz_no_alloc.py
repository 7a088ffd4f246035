"""thorin_rad_solve_noalloc, the energy solve that allocates nothing, against
thorin_rad_solve on the same assembled systems: the same temperature bit for
bit, the same sweeps, rounds and residuals, for both methods.  Two calls in a
row on one GPU, the second on a smaller grid (the shared device array keeps
the first call's contents); a grid above THORIN_NOALLOC_MAX_CELLS and a warm
start are refused."""
import numpy as np
import pytest

from fargo_thorin_amd import Disk, Params, ThorinError, WarmStart, configs
from fargo_thorin_amd.seed import sor_disk

pytestmark = pytest.mark.gpu


def _disk(nr, ns):
    print(f"[gpu-noalloc] {nr}x{ns}", flush=True)
    d = Disk(Params(**configs.with_grid(configs.SOR_2048x3072, nr, ns), OutputDir="/nonexistent-thorin-out/"))
    disk = sor_disk(nr, ns)
    d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
    d.rad_system(0.0)
    return d


def _same(a, b):
    for k in ("sweeps", "rounds", "converged", "resid0", "resid"):
        assert a[k] == b[k], (k, a[k], b[k])


@pytest.mark.parametrize("method,kw", [(1, {"inner_sweeps": 12}), (1, {"inner_sweeps": 7, "sweep_block": 1}),
                                       (0, {})])
def test_noalloc_solve_equals_allocating_solve(method, kw):
    for nr, ns in ((384, 576), (130, 256)):   # the second call on a smaller grid
        d = _disk(nr, ns)
        ref = d.rad_solve(method, tol=1e-10, omega=1.8, **kw)
        t_ref = d.field("temper")
        got = d.rad_solve(method, tol=1e-10, omega=1.8, noalloc=True, **kw)
        t_got = d.field("temper")
        _same(ref, got)
        assert got["converged"]
        assert np.array_equal(t_got, t_ref), f"{nr}x{ns}: max |diff| {np.max(np.abs(t_got - t_ref)):.3e}"
        d.close()


def test_noalloc_solve_twice_on_one_context():
    d = _disk(256, 384)
    a = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=10, noalloc=True)
    ta = d.field("temper")
    b = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=10, noalloc=True)
    _same(a, b)
    assert np.array_equal(d.field("temper"), ta)


def test_noalloc_refuses_large_grid_and_warm_start():
    d = _disk(1100, 1500)   # (1100 + 1) * 1500 = 1651500 cells > THORIN_NOALLOC_MAX_CELLS
    with pytest.raises(ThorinError, match="THORIN_NOALLOC_MAX_CELLS"):
        d.rad_solve(1, tol=1e-10, omega=1.8, noalloc=True)
    r = d.rad_solve(1, tol=1e-10, omega=1.8)
    assert r["converged"]
    small = _disk(64, 96)
    with pytest.raises(ThorinError, match="no warm start"):
        small.rad_solve(1, tol=1e-10, omega=1.8, noalloc=True, warm=WarmStart(3))
