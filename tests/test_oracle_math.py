"""Closed forms and invariants that pin the oracle's pointwise laws.

These checks do not restate the oracle's formulas.  Each one tests a property
the mathematics fixes, so that a dropped term, a wrong exponent or a swapped
coefficient fails:

  * opacity laws (reference src_main/Opacity.c): every regime of the
    Bell & Lin (1994) and Lin & Papaloizou (1985) tables is a power law
    kappa_i rho^a_i T^b_i with the published (kappa_i, a_i, b_i), so deep
    inside a regime the local logarithmic slopes d ln kappa / d ln T and
    d ln kappa / d ln rho equal (b_i, a_i), and the regime borders are where
    neighbouring power laws are equal, so the law is continuous there.  The
    Zhu et al. (2012) fit is piecewise linear in (log T, log P) and continuous
    at the borders its authors corrected (Opacity.c:141-143).
  * POWERLAW (kappa = k0 T^2): exact scaling kappa(2T) = 4 kappa(T), and on a
    disk colder than the ice line the whole hydro step with k0 = 2e-4 gives
    the same bits as the reference's LP85, whose ice-grain branch is
    2e-4 T^2 (Opacity.c:203-205); that run is a reference golden
    (cold_lp85_48x96).
  * flux limiter (EnergySources.c:684-696, Kley 1989): lambda(0) = 1/3 (the
    diffusion limit), lambda R -> 1 (free streaming), continuity at R = 2,
    monotone decrease, and within 30% of the Levermore & Pomraning (1981)
    limiter (2 + R)/(6 + 3R + R^2) that it approximates (27% at most).
"""
import math

import numpy as np
import pytest

from oracle import OracleDisk, lib
from tests.helpers import Replay, case, golden, params_for, planets_for

BL94, LP85, ZHU12, CONST, POWERLAW = 0, 1, 2, 3, 4


def kappa(law, rho, T, k0=0.0):
    return lib().oracle_opacity_law(law, float(rho), float(T), float(k0))


def slopes(law, rho, T, h=1e-4):
    """Central log-log derivatives (d ln k/d ln T, d ln k/d ln rho)."""
    kt = (math.log(kappa(law, rho, T * math.exp(h))) - math.log(kappa(law, rho, T * math.exp(-h)))) / (2 * h)
    kr = (math.log(kappa(law, rho * math.exp(h), T)) - math.log(kappa(law, rho * math.exp(-h), T))) / (2 * h)
    return kt, kr


# (T in K, rho in g/cm^3, published exponents (b, a)) deep inside each regime
# Bell & Lin (1994) Table 3 as Keith & Wardle (2014) Table 1 gives it
BL94_POINTS = [
    (40.0, 1e-9, (2.0, 0.0)),         # ice grains
    (600.0, 1e-9, (0.5, 0.0)),        # metal grains
    (1200.0, 1e-9, (-24.0, 1.0)),     # metal evaporation
    (2500.0, 1e-9, (3.0, 2.0 / 3.0)),   # molecules
    (5000.0, 1e-9, (10.0, 1.0 / 3.0)),  # H- scattering
    (2.0e4, 1e-9, (-2.5, 1.0)),       # bound-free / free-free
    (1.0e7, 1e-9, (0.0, 0.0)),        # electron scattering
]
# (the ice-evaporation regime of BL94 is blended with its neighbours by the
# smooth low-temperature combination, Opacity.c:47-55, and has no pure slope)

# Lin & Papaloizou (1985) as Mueller & Kley (2012) Table 1 gives it
LP85_POINTS = [
    (60.0, 1e-9, (2.0, 0.0)),         # ice grains
    (185.0, 1e-9, (-7.0, 0.0)),       # ice evaporation
    (600.0, 1e-9, (1.0, 0.0)),        # silicate grains
    (2000.0, 1e-9, (-9.0, 2.0 / 3.0)),  # silicate evaporation
    (3700.0, 1e-9, (3.0, 2.0 / 3.0)),   # molecules
    (5000.0, 1e-9, (10.0, 1.0 / 3.0)),  # H- scattering
    (3.0e4, 1e-9, (-2.5, 1.0)),       # bound-free / free-free
]


@pytest.mark.parametrize("law,points", [(BL94, BL94_POINTS), (LP85, LP85_POINTS)], ids=["BL94", "LP85"])
def test_opacity_regimes_have_published_exponents(law, points):
    for T, rho, (b, a) in points:
        kt, kr = slopes(law, rho, T)
        assert abs(kt - b) < 2e-3 and abs(kr - a) < 2e-3, f"T={T}: slopes {kt:.4f},{kr:.4f} vs {b},{a}"


def test_opacity_regime_coefficients():
    """The prefactors: in the ice-grain regimes both laws are 2e-4 T^2; LP85's
    silicate grains 5e-3 T; BL94's electron scattering 0.348 cm^2/g."""
    for T in (20.0, 30.0, 40.0):
        assert kappa(LP85, 1e-9, T) == pytest.approx(2e-4 * T * T, rel=1e-15)
        # BL94 blends in the ice evaporation law, (k1/k2)^2 < 1e-11 below 40 K
        assert kappa(BL94, 1e-9, T) == pytest.approx(2e-4 * T * T, rel=1e-10)
    assert kappa(LP85, 1e-9, 800.0) == pytest.approx(5e-3 * 800.0, rel=1e-15)
    assert kappa(BL94, 1e-9, 1e7) == 0.348


def _borders(law, rho, lo, hi, n=20000):
    """Temperatures where the local T-slope jumps (regime borders)."""
    T = np.exp(np.linspace(math.log(lo), math.log(hi), n))
    k = np.array([math.log(kappa(law, rho, t)) for t in T])
    s = np.diff(k) / np.diff(np.log(T))
    jumps = np.where(np.abs(np.diff(s)) > 0.5)[0]
    return T, k, jumps


@pytest.mark.parametrize("law", [BL94, LP85], ids=["BL94", "LP85"])
@pytest.mark.parametrize("rho", [1e-11, 1e-9, 1e-7])
def test_opacity_continuous_at_regime_borders(law, rho):
    """Borders are where neighbouring power laws are equal: across each
    border the jump of ln kappa is no larger than the slope times the step.
    Up to 6000 K for BL94 at rho = 1e-11: above it the H- scattering regime
    ends before it begins at that density (the borders are out of order, and
    the reference's law jumps by 0.31 in ln kappa at 8076 K, as the
    reference notes for densities outside the fit's range, Opacity.c:22-23)."""
    hi = 6000.0 if (law == BL94 and rho < 1e-10) else 1e6
    T, k, jumps = _borders(law, rho, 20.0, hi)
    assert len(jumps) >= 5
    dlnT = math.log(T[1] / T[0])
    for j in jumps:
        assert abs(k[j + 2] - k[j + 1]) < 30 * dlnT, f"jump at T={T[j + 1]:.1f}: {k[j + 2] - k[j + 1]:.3e}"


def test_zhu12_continuous_at_corrected_borders():
    """Zhu et al. (2012): continuous in log T across the borders 3->4, 4->5,
    5->6 and 6->7 that the reference corrected (Opacity.c:141-143)."""
    for rho in (1e-11, 1e-9, 1e-7):
        T, k, jumps = _borders(ZHU12, rho, 300.0, 3000.0, n=40000)
        dlnT = math.log(T[1] / T[0])
        assert len(jumps) >= 3
        for j in jumps:
            assert abs(k[j + 2] - k[j + 1]) < 60 * dlnT, f"rho={rho} T={T[j + 1]:.1f}"


def test_powerlaw_scaling_exact():
    for k0 in (2e-4, 3.7e-3):
        for T in (20.0, 150.0, 1999.0):
            assert kappa(POWERLAW, 1e-9, 2 * T, k0) == 4 * kappa(POWERLAW, 1e-9, T, k0)
            assert kappa(POWERLAW, 1e-5, T, k0) == kappa(POWERLAW, 1e-12, T, k0)
        assert kappa(POWERLAW, 1e-9, 1.0, k0) == k0


def test_powerlaw_equals_lp85_ice_branch_bitwise():
    """Below the ice line (166.8 K) LP85 is 2e-4 T^2: POWERLAW with k0 = 2e-4
    gives the same bits."""
    for T in np.linspace(10.0, 166.0, 157):
        assert kappa(POWERLAW, 1e-9, T, 2e-4) == kappa(LP85, 1e-9, T)


def test_powerlaw_hydro_step_equals_reference_lp85_on_cold_disk():
    """The whole step with OpacityTable POWERLAW, PowerLawOpacity 2e-4 on the
    cold disk (64-132 K) equals the reference's LP85 run bit for bit."""
    name = "cold_lp85_48x96"
    cfg, planets, _ = case(name)
    g = golden(name)
    o = OracleDisk(params_for(cfg, OpacityTable="POWERLAW", PowerLawOpacity="2e-4"))
    o.init(g["init_dens"], g["init_vrad"], g["init_vtheta"], g["init_temper"])
    sysp = planets_for(planets, cfg)
    rep = Replay(g["tqwk0"])
    niters = []
    ntot, ninterm = int(g["ntot"]), int(g["ninterm"])
    for i in range(ntot):
        lg = o.algogas(sysp, rep.cb)
        niters += [lg.sor_niter[k] for k in range(lg.nsub)]
        if (i + 1) % ninterm == 0:
            n = (i + 1) // ninterm
            for f in ("dens", "vrad", "vtheta", "temper", "energy", "opacity"):
                assert np.array_equal(o.field(f), g[f"{f}{n}"]), f"{f}{n}"
    assert niters == list(g["sor_niter"][:len(niters)])


def test_flux_limiter_limits():
    fl = lib().oracle_flux_limiter
    assert fl(0.0) == pytest.approx(1.0 / 3.0, rel=1e-15)
    assert fl(2.0) == pytest.approx(0.2, rel=1e-15)
    assert fl(2.0 + 1e-12) == pytest.approx(0.2, rel=1e-10)
    for R in (1e3, 1e6, 1e9):
        assert fl(R) * R == pytest.approx(1.0, rel=10.0 / math.sqrt(R))
    R = np.exp(np.linspace(math.log(1e-3), math.log(1e4), 2000))
    lam = np.array([fl(r) for r in R])
    assert np.all(np.diff(lam) < 0)
    lp = (2.0 + R) / (6.0 + 3.0 * R + R * R)
    assert np.max(np.abs(lam / lp - 1.0)) < 0.3
