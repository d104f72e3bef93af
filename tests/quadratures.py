"""Exact answers for a few photon orbits, from quadratures of the Kerr orbit equations with
no integrator in the loop: the oracle tests and the GPU tests hold the traced rays to them.
E = 1 throughout; L and Q are the Bardeen launch's (src/eht.cpp:262-271)."""
import math

import numpy as np
from scipy import integrate, optimize


def schwarzschild_sweep(b, r_cam):
    """Angle swept by an equatorial a = 0 photon of impact parameter b from r_cam in to
    periapsis and out to infinity: integral of du / sqrt(1/b^2 - u^2 + 2u^3), with
    u = u0 - t^2 removing the inverse square root at the turning point u0.
    (inward, outward, periapsis radius)"""
    roots = np.roots([2.0, -1.0, 0.0, 1.0 / (b * b)])
    u0 = min(x.real for x in roots if abs(x.imag) < 1e-12 and x.real > 0)

    def g(u):  # B(u) = (u0 - u) g(u)
        return (u + u0) - 2.0 * (u * u + u * u0 + u0 * u0)

    f = lambda t: 2.0 / math.sqrt(g(u0 - t * t))  # noqa: E731
    uc = 1.0 / r_cam
    inward = integrate.quad(f, 0.0, math.sqrt(u0 - uc), epsabs=1e-14, epsrel=1e-13, limit=200)[0]
    outward = integrate.quad(f, 0.0, math.sqrt(u0), epsabs=1e-14, epsrel=1e-13, limit=200)[0]
    return inward, outward, 1.0 / u0


def schwarzschild_pole_disk_hit(b, r_cam):
    """a = 0, camera on the equator, L = 0, impact parameter b: the ray crosses the pole and
    meets the equatorial plane on the far side where the swept angle reaches pi, just past
    periapsis.  The radius there."""
    inward, outward, r_min = schwarzschild_sweep(b, r_cam)
    assert inward + outward > math.pi
    u0 = 1.0 / r_min

    def g(u):
        return (u + u0) - 2.0 * (u * u + u * u0 + u0 * u0)

    f = lambda t: 2.0 / math.sqrt(g(u0 - t * t))  # noqa: E731
    need = math.pi - inward
    assert need > 0
    # the swept angle from periapsis out to u = u0 - t^2 is the integral of f over [0, t]
    t_hit = optimize.brentq(lambda t: integrate.quad(f, 0.0, t, epsabs=1e-14, epsrel=1e-13)[0] - need, 0.0,
                            math.sqrt(u0), xtol=1e-15)
    return 1.0 / (u0 - t_hit * t_hit)


def kerr_equatorial_sweep(a, L, r_cam):
    """Angle swept by an equatorial (Q = 0, E = 1) photon of angular momentum L around a
    hole of spin a, from r_cam in to periapsis and out to infinity.  With u = 1/r the
    orbit equations give dphi = F(u) du / sqrt(U(u)):
        U(u) = 1 + (a^2 - L^2) u^2 + 2 (L - a)^2 u^3          (u^4 R(1/u))
        F(u) = L - a + a (1 + a (a - L) u^2) / (1 - 2 u + a^2 u^2)   (L - aE + a P/Delta)
    and u = u0 - t^2 removes the inverse square root at periapsis u0, with
    U(u) = (u0 - u) V(u), V(u) = -(c3 (u^2 + u u0 + u0^2) + c2 (u + u0)).
    (inward, outward, periapsis radius)"""
    c3, c2 = 2.0 * (L - a) ** 2, a * a - L * L
    roots = np.roots([c3, c2, 0.0, 1.0])
    u0 = min(x.real for x in roots if abs(x.imag) < 1e-12 and x.real > 0)

    def f(t):
        u = u0 - t * t
        V = -(c3 * (u * u + u * u0 + u0 * u0) + c2 * (u + u0))
        F = L - a + a * (1.0 + a * (a - L) * u * u) / (1.0 - 2.0 * u + a * a * u * u)
        return 2.0 * F / math.sqrt(V)

    uc = 1.0 / r_cam
    inward = integrate.quad(f, 0.0, math.sqrt(u0 - uc), epsabs=1e-14, epsrel=1e-13, limit=400)[0]
    outward = integrate.quad(f, 0.0, math.sqrt(u0), epsabs=1e-14, epsrel=1e-13, limit=400)[0]
    return inward, outward, 1.0 / u0


def kerr_equatorial_phi(a, alpha, r_cam):
    """Sky phi of the edge-on Bardeen ray (alpha, beta = 0) of a camera at phi = 0: the
    photon has L = -alpha and is traced backward, so phi runs the other way."""
    inward, outward, _ = kerr_equatorial_sweep(a, -alpha, r_cam)
    return (-(inward + outward)) % (2 * math.pi)


def kerr_disk_hit(a, r_cam, inc, alpha, beta):
    """Bardeen ray (alpha, beta < 0) of a camera at inclination inc < pi/2 that runs down to
    the equator with neither r nor theta turning: at the crossing the Mino times agree,
    int_inc^(pi/2) dtheta / sqrt(Theta) = int_(r_hit)^(r_cam) dr / sqrt(R).  r_hit from that
    equation (quadrature + root)."""
    Lz = -alpha * math.sin(inc)
    Q = beta * beta + (alpha * alpha - a * a) * math.cos(inc) ** 2

    def theta_pot(t):
        return Q - math.cos(t) ** 2 * (Lz * Lz / math.sin(t) ** 2 - a * a)

    def radial_pot(r):
        return (r * r + a * a - a * Lz) ** 2 - (r * r - 2.0 * r + a * a) * (Q + (Lz - a) ** 2)

    assert beta < 0 and min(theta_pot(t) for t in np.linspace(inc, math.pi / 2, 2001)) > 0
    lam = integrate.quad(lambda t: 1.0 / math.sqrt(theta_pot(t)), inc, math.pi / 2, epsabs=1e-15, epsrel=1e-13,
                         limit=400)[0]
    roots = np.roots([1.0, 0.0, a * a - Lz * Lz - Q, 2.0 * (Q + (Lz - a) ** 2), -a * a * Q])
    r_turn = max([x.real for x in roots if abs(x.imag) < 1e-10 and x.real < r_cam] +
                 [1.0 + math.sqrt(1.0 - a * a)])

    def mino_gap(r):
        return integrate.quad(lambda x: 1.0 / math.sqrt(radial_pot(x)), r, r_cam, epsabs=1e-15, epsrel=1e-13,
                              limit=400)[0] - lam

    lo = r_turn + 1e-3
    assert mino_gap(lo) > 0  # the equator comes before the radial turning point
    return optimize.brentq(mino_gap, lo, r_cam, xtol=1e-14, rtol=1e-15)


def keplerian_g(a, r, Lz):
    """g = sqrt(-(g_tt + 2 g_tphi Omega + g_phiphi Omega^2)) / (1 - L Omega) for prograde
    Keplerian gas, Omega = 1 / (r^1.5 + a), in the Boyer-Lindquist equatorial metric."""
    om = 1.0 / (r ** 1.5 + a)
    return math.sqrt((1 - 2 / r) + 4 * a / r * om - (r * r + a * a + 2 * a * a / r) * om * om) / (1 - Lz * om)
