"""The oracle's massive particles (ko_particle_advance, ko_particle_from_zamo): debris on
Kerr geodesics advanced by a span of Kerr-Schild coordinate time.  Pinned by

  - golden vectors from the reference (tests/golden/particle_vectors.npz, from
    tests/golden/make_particle_golden.py): kerr::launchFromZamo (bit for bit) and
    trajectory::predict's orbits (bound, inclined, over the poles, plunging, escaping,
    Schwarzschild, near-extremal): fates, positions after 200-400 M, radial ranges;
  - closed forms: a circular equatorial orbit keeps its radius and turns at
    Omega = 1 / (r^1.5 + a); radial infall from rest reaches r at the Kerr-Schild time of
    the exact integral, through the horizon; an eccentric Schwarzschild orbit returns to
    periapsis after the radial period of the exact integral, having turned by its exact
    periapsis advance; the state stays on vr^2 = R(r), vmu^2 = M(mu).
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O

G = np.load(__file__.rsplit("/", 1)[0] + "/golden/particle_vectors.npz")
L = O.lib()
OPTS = O.make_opts()


def cartesian(a, st):
    th, ph = math.acos(max(-1.0, min(1.0, st[1]))), st[2]
    return np.array([(st[0] * math.cos(ph) - a * math.sin(ph)) * math.sin(th),
                     (st[0] * math.sin(ph) + a * math.cos(ph)) * math.sin(th), st[0] * math.cos(th)])


def test_launch_from_zamo_matches_reference():
    for (a, r, th, v0, v1, v2), ref in zip(G["zamo_in"], G["zamo_out"]):
        got = O.particle_from_zamo(a, r, th, [v0, v1, v2])
        assert np.array_equal(got[:4], ref[:4])
        assert got[4] == -math.sin(th) * ref[4]  # vmu = -sin(theta) vtheta


def test_orbits_match_reference_predict():
    """trajectory::predict (trajectory.cpp:121-172): the same fate; bound orbits end at the
    same point after 200-400 M (the reference's RK4 in coordinate time agrees to ~1e-9 M)
    and cover the same radial range."""
    fates = {0: O.PARTICLE_MOVING, 2: O.PARTICLE_PLUNGED, 3: O.PARTICLE_ESCAPED}
    for (a, r, th, v0, v1, v2, tmax, esc), ref in zip(G["predict_in"], G["predict_out"]):
        E, Lz, Q, vr, vmu = O.particle_from_zamo(a, r, th, [v0, v1, v2])
        rp = L.ko_horizon_radius(a)
        st = np.array([r, math.cos(th), 0.0, vr, vmu])
        rmin = rmax = r
        t, fate = 0.0, O.PARTICLE_MOVING
        while t < tmax - 1e-9 and fate == O.PARTICLE_MOVING:
            dt = min(0.5, tmax - t)
            fate, st, _ = O.particle_advance(a, E, Lz, Q, st, dt, rp + 0.02, esc, OPTS)
            t += dt
            if fate == O.PARTICLE_MOVING:
                rmin, rmax = min(rmin, st[0]), max(rmax, st[0])
        assert fate == fates[int(ref[0])], (a, r, fate, ref[0])
        if fate == O.PARTICLE_MOVING:
            x = cartesian(a, st)
            assert np.linalg.norm(x - ref[1:4]) < 1e-7 * np.linalg.norm(ref[1:4])
            assert rmin == pytest.approx(ref[4], rel=1e-4) and rmax == pytest.approx(ref[5], rel=1e-4)
        else:  # within a step of the reference's (4 M long at 400 M), and of our 0.5 M calls
            assert t == pytest.approx(ref[8], abs=5.0)


def circular_el(a, r):
    sr = math.sqrt(r)
    den = r ** 0.75 * math.sqrt(r * sr - 3 * sr + 2 * a)
    return (r * sr - 2 * sr + a) / den, (r * r - 2 * a * sr + a * a) / den


@pytest.mark.parametrize("a,r", [(0.9, 8.0), (0.0, 10.0), (0.998, 3.0)])
def test_circular_equatorial_orbit(a, r):
    E, Lz = circular_el(a, r)
    st = np.array([r, 0.0, 0.0, 0.0, 0.0])
    T = 2000.0
    fate, st, steps = O.particle_advance(a, E, Lz, 0.0, st, T, 1.0, 1e4, OPTS)
    assert fate == O.PARTICLE_MOVING
    assert st[0] == pytest.approx(r, rel=1e-8) and abs(st[1]) < 1e-12
    omega = 1.0 / (r ** 1.5 + a)
    dphi = (st[2] - omega * T + math.pi) % (2 * math.pi) - math.pi
    assert abs(dphi) < 1e-8 * omega * T


def infall_t(r0, r1):
    """Kerr-Schild time for a = 0 radial infall from rest at infinity from r0 to r1:
    dt_BL/dr = -r^1.5 / (sqrt(2) (r - 2)), t_KS = t_BL + 2 ln|r/2 - 1|."""
    f = lambda r: (2.0 - r ** 1.5 / math.sqrt(2.0)) / (r - 2.0) if abs(r - 2.0) > 1e-9 else -1.5
    return integrate.quad(f, r0, r1, epsabs=1e-13, epsrel=1e-13, limit=200)[0]


@pytest.mark.parametrize("r1", [6.0, 2.5, 1.5, 0.6])
def test_radial_infall_time_through_the_horizon(r1):
    a, r0 = 0.0, 20.0
    vr = -math.sqrt(2.0 * r0 ** 3)  # R(r) = r^4 - (r^2 - 2r) r^2 = 2 r^3 for E = 1, L = Q = 0
    dt = infall_t(r0, r1)
    fate, st, _ = O.particle_advance(a, 1.0, 0.0, 0.0, np.array([r0, 0.3, 1.0, vr, 0.0]), dt, 0.1, 1e4, OPTS)
    assert fate == O.PARTICLE_MOVING
    assert st[0] == pytest.approx(r1, rel=5e-9)
    assert st[1] == 0.3 and st[2] == 1.0  # radial: theta and phi stay


def test_schwarzschild_radial_period_and_periapsis_advance():
    """r_p = 8, r_a = 20: E^2 and L^2 solve R(r_p) = R(r_a) = 0; the radial period and the
    advance are 2 * integral over [r_p, r_a] of dt/dr and dphi/dr (r = p / (1 + e cos chi))."""
    rp_, ra_ = 8.0, 20.0
    A = np.array([[rp_ ** 4, -(rp_ ** 2 - 2 * rp_)], [ra_ ** 4, -(ra_ ** 2 - 2 * ra_)]])
    b = np.array([(rp_ ** 2 - 2 * rp_) * rp_ ** 2, (ra_ ** 2 - 2 * ra_) * ra_ ** 2])
    E2, L2 = np.linalg.solve(A, b)
    E, Lz = math.sqrt(E2), math.sqrt(L2)
    p, e = 2 * rp_ * ra_ / (rp_ + ra_), (ra_ - rp_) / (ra_ + rp_)

    def dr_dchi(chi):
        return p * e * math.sin(chi) / (1 + e * math.cos(chi)) ** 2

    def rates(chi):
        r = p / (1 + e * math.cos(chi))
        R = E2 * r ** 4 - (r * r - 2 * r) * (r * r + L2)
        sq = math.sqrt(max(R, 1e-300))
        # dt_BL/dlambda = E r^4 / (r^2 - 2r) for a = 0; dphi/dlambda = L
        return r, E * r ** 4 / (r * r - 2 * r) / sq, Lz / sq

    Tr = 2 * integrate.quad(lambda c: rates(c)[1] * dr_dchi(c), 1e-12, math.pi - 1e-12, epsabs=1e-12,
                            epsrel=1e-13, limit=400)[0]
    dphi = 2 * integrate.quad(lambda c: rates(c)[2] * dr_dchi(c), 1e-12, math.pi - 1e-12, epsabs=1e-13,
                              epsrel=1e-13, limit=400)[0]
    assert dphi > 2 * math.pi  # the periapsis advances
    st = np.array([rp_, 0.0, 0.0, 0.0, 0.0])
    fate, st, _ = O.particle_advance(0.0, E, Lz, 0.0, st, Tr, 2.02, 1e4, OPTS)
    assert fate == O.PARTICLE_MOVING
    assert st[0] == pytest.approx(rp_, rel=2e-8)
    assert abs((st[2] - dphi + math.pi) % (2 * math.pi) - math.pi) < 2e-7


def test_polar_orbit_passes_the_poles_smoothly_and_stays_on_the_constraints():
    """L = 0 with enough Q to reach the poles: phi gains pi at each pass, so the Cartesian path
    is continuous (no jump between samples 0.25 M apart); one call over 150 M ends where
    600 calls of 0.25 M end; vr^2 = R and vmu^2 = M hold throughout."""
    a, r, th = 0.5, 10.0, 1.0
    E, Lz, Q, vr, vmu = O.particle_from_zamo(a, r, th, [0.0, 0.35, 0.0])
    assert Lz == 0.0 and Q > 0
    st = np.array([r, math.cos(th), 0.0, vr, vmu])
    x0 = cartesian(a, st)
    passes = 0
    for _ in range(600):
        prev_mu = st[1]
        fate, st, _ = O.particle_advance(a, E, Lz, Q, st, 0.25, 1.0, 1e4, OPTS)
        x1 = cartesian(a, st)
        assert np.linalg.norm(x1 - x0) < 0.25 * 1.0 + 1e-9  # |dx/dt| < 1
        x0 = x1
        passes += abs(st[1]) > 0.999 or (prev_mu > 0.999 and st[1] < prev_mu)
        r_, m_ = st[0], st[1]
        delta = r_ * r_ - 2 * r_ + a * a
        P = E * (r_ * r_ + a * a)
        R = P * P - delta * (r_ * r_ + (Lz - a * E) ** 2 + Q)
        A2E2 = a * a * (E * E - 1)
        M = Q - (Q + Lz * Lz - A2E2) * m_ ** 2 - A2E2 * m_ ** 4
        assert st[3] ** 2 == pytest.approx(R, rel=1e-6, abs=1e-8 * r_ ** 4)
        assert st[4] ** 2 == pytest.approx(M, abs=1e-8 * (Q + 1))
    assert passes >= 1
    fate, one, _ = O.particle_advance(a, E, Lz, Q, np.array([r, math.cos(th), 0.0, vr, vmu]), 150.0, 1.0, 1e4, OPTS)
    assert np.linalg.norm(cartesian(a, one) - x0) < 1e-7


def test_carried_step_size_saves_attempts_and_keeps_the_orbit():
    """The step size handed from one update to the next: the orbit is the same to the
    tolerance, and the updates after the first take fewer attempts than cold starts."""
    a, r, th = 0.9, 12.0, 1.2
    E, Lz, Q, vr, vmu = O.particle_from_zamo(a, r, th, [0.1, 0.2, 0.3])
    cold = warm = np.array([r, math.cos(th), 0.0, vr, vmu])
    h, n_cold, n_warm = 0.0, 0, 0
    for _ in range(30):
        _, cold, s1 = O.particle_advance(a, E, Lz, Q, cold, 10.0, 1.45, 1e4, OPTS)
        _, warm, s2, h = O.particle_advance(a, E, Lz, Q, warm, 10.0, 1.45, 1e4, OPTS, h=h)
        n_cold += s1
        n_warm += s2
        assert h > 0.0
    assert n_warm < 0.8 * n_cold
    assert np.linalg.norm(cartesian(a, cold) - cartesian(a, warm)) < 1e-7
