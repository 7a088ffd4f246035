"""The host N-body part on the CPU: the library's integrator against the
closed-form Kepler orbit, the orbital elements of nbody.orbits.dat against
the two-body relations and against the rows the reference wrote (REBOUND's
elements of the golden runs' planets)."""
import ctypes as C

import numpy as np
import pytest

from fargo_thorin_amd import make_planets
from fargo_thorin_amd._capi import host_nbody_fn, last_error, lib
from tests.helpers import golden


def _state(mu, a, e, inc, Om, om, f):
    """Heliocentric position and velocity of an elliptic orbit: the
    perifocal frame rotated by Omega, inc, omega."""
    p = a * (1.0 - e * e)
    r = p / (1.0 + e * np.cos(f))
    pos_pf = np.array([r * np.cos(f), r * np.sin(f), 0.0])
    vel_pf = np.sqrt(mu / p) * np.array([-np.sin(f), e + np.cos(f), 0.0])

    def rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1.0]])

    def rx(t):
        return np.array([[1.0, 0, 0], [0, np.cos(t), -np.sin(t)], [0, np.sin(t), np.cos(t)]])

    R = rz(Om) @ rx(inc) @ rz(om)
    return R @ pos_pf, R @ vel_pf


def _planet(m, pos, vel):
    s = make_planets([m], [1.0])
    s.mass[0] = m                    # make_planets rounds masses to float, as the planet file is read
    s.x[0], s.y[0], s.z[0] = pos
    s.vx[0], s.vy[0], s.vz[0] = vel
    return s


def _elements(s, k=0, mp=1.0):
    el = (C.c_double * 6)()
    assert lib().thorin_orbit_elements(C.byref(s), k, mp, el) == 0, last_error()
    return np.array(el[:])


def _wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


@pytest.mark.parametrize("a,e,inc,Om,om,f", [
    (1.0, 0.3, 0.2, 0.7, 1.1, 2.0),
    (5.2, 0.05, 0.4, -2.0, 0.3, -1.0),
    (0.7, 0.6, 1.2, 2.9, -2.5, 0.1),
    (2.0, 0.1, 2.5, 1.0, 1.0, 1.0),          # retrograde
])
def test_elements_recover_the_orbit(a, e, inc, Om, om, f):
    m = 3e-5
    pos, vel = _state(1.0 + m, a, e, inc, Om, om, f)
    el = _elements(_planet(m, pos, vel))
    assert abs(el[0] - a) <= 1e-12 * a and abs(el[1] - e) <= 1e-13
    assert abs(el[2] - inc) <= 1e-13
    for got, want in zip(el[3:], (Om, om, f)):
        assert abs(_wrap(got - want)) <= 1e-12
    # angles in (-pi, pi]
    assert np.all(np.abs(el[3:]) <= np.pi)


def test_planar_orbit_conventions():
    """In the plane the node is undefined: Omega = 0, omega is the longitude
    of the pericentre, f the planet's longitude minus it."""
    m = 1e-4
    pos, vel = _state(1.0 + m, 1.3, 0.2, 0.0, 0.0, 0.9, -0.4)
    el = _elements(_planet(m, pos, vel))
    assert el[2] == 0.0 and el[3] == 0.0
    assert abs(el[4] - 0.9) <= 1e-12 and abs(el[5] + 0.4) <= 1e-12
    # a circular orbit: no pericentre, omega = 0 and f is the longitude
    el = _elements(_planet(0.0, np.array([np.cos(2.5), np.sin(2.5), 0.0]),
                           np.array([-np.sin(2.5), np.cos(2.5), 0.0])))
    assert el[1] < 1e-15 and el[4] == 0.0 and abs(el[5] - 2.5) <= 1e-15


def test_elements_refuse_a_planet_at_the_primary():
    s = _planet(1e-5, np.zeros(3), np.array([0.0, 1.0, 0.0]))
    el = (C.c_double * 6)()
    assert lib().thorin_orbit_elements(C.byref(s), 0, 1.0, el) != 0
    assert lib().thorin_orbit_elements(C.byref(s), 1, 1.0, el) != 0


@pytest.mark.parametrize("name", ["flock_cavity_48x96", "hill_q1e-4_48x96", "pebbles_q3e-5_64x128",
                                  "pebbles_1024x1536_100"])
def test_elements_of_the_reference_rows(name):
    """Each nbody.orbits.dat row of the reference (REBOUND's elements of its
    planet, %.12g) against the elements of the row's own position and
    velocity.  The rows carry 12 digits, so a and e agree to ~1e-11; of the
    angles, the longitude Omega + omega + f of the planet is well defined at
    any eccentricity and omega alone only where e is far above 1e-11."""
    orb = golden(name)["nbody_orbits"]
    mass = orb[0, 8]                      # the set-up row has MassTaper 1
    for row in orb:
        # the mass column is the (accreting) mass times MassTaper, 0 in the
        # first DT and 1 after it in these runs
        mass = row[8] if row[8] > 0 else mass
        el = _elements(_planet(mass, row[12:15], row[15:18]))
        assert abs(el[0] - row[2]) <= 2e-11 * row[2], row[:2]
        assert abs(el[1] - row[3]) <= 2e-11, row[:2]
        assert abs(el[2] - row[4]) <= 1e-11 and abs(el[3] - row[5]) <= 1e-11
        assert abs(_wrap(el[3] + el[4] + el[5] - row[5] - row[6] - row[7])) <= 1e-10, row[:2]
        if row[3] > 1e-4:
            assert abs(_wrap(el[4] - row[6])) <= 1e-6, row[:2]


def test_planet_radius_column_from_planetary_density():
    """The radius column: the sphere of the planet's mass at PlanetaryDensity
    1 g/cm^3 in code units (solar mass, au)."""
    for name in ("flock_cavity_48x96", "hill_q1e-4_48x96"):
        row = golden(name)["nbody_orbits"][0]
        msol_g, au_cm = 1.98855e33, 1.495978707e13
        rho = 1.0 / (msol_g / au_cm**3)
        r = (3.0 * row[8] / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
        assert abs(r - row[11]) <= 1e-11 * r


def _kepler(mu, a, e, inc, Om, om, f0, t):
    """Closed-form position after time t: mean anomaly advanced at
    n = sqrt(mu / a^3), Kepler's equation E - e sin E = M by Newton."""
    E0 = 2.0 * np.arctan(np.sqrt((1 - e) / (1 + e)) * np.tan(f0 / 2.0))
    M = E0 - e * np.sin(E0) + np.sqrt(mu / a**3) * t
    E = M
    for _ in range(50):
        E -= (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
    f = 2.0 * np.arctan2(np.sqrt(1 + e) * np.sin(E / 2.0), np.sqrt(1 - e) * np.cos(E / 2.0))
    return _state(mu, a, e, inc, Om, om, f)


@pytest.mark.parametrize("e,inc", [(0.0, 0.0), (0.3, 0.2), (0.7, 1.0)])
def test_nbody_step_follows_the_kepler_orbit(e, inc):
    """One planet, many substeps of the library's N-body step (the
    integrator of the standalone driver) against the closed form over two
    orbits."""
    m, a = 3e-5, 1.0
    mu = 1.0 + m
    pos, vel = _state(mu, a, e, inc, 0.4, 0.8, 0.3)
    s = _planet(m, pos, vel)
    nb = host_nbody_fn()
    dt, nsteps = 0.02, int(round(2 * 2 * np.pi / 0.02))
    for k in range(nsteps):
        assert nb(None, C.byref(s), k * dt, dt, 0.0) == 0
    want_p, want_v = _kepler(mu, a, e, inc, 0.4, 0.8, 0.3, nsteps * dt)
    got_p = np.array([s.x[0], s.y[0], s.z[0]])
    got_v = np.array([s.vx[0], s.vy[0], s.vz[0]])
    assert np.max(np.abs(got_p - want_p)) <= 1e-11, np.max(np.abs(got_p - want_p))
    assert np.max(np.abs(got_v - want_v)) <= 1e-11 * (1 + e) / (1 - e)
