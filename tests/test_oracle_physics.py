"""The FP64 oracle against what the mathematics fixes: the reference's own theta-form
equations (chain rule), finite differences of the potentials, conserved quantities,
exact Schwarzschild and Kerr quadratures (quadratures.py), the critical curve and the
reference's known answers (test_kerr.cpp testShadow, testNearAxisPhotons, testEht)."""
import math

import numpy as np
import pytest
import quadratures as Qd
from scipy import optimize

from oracle import oracle as O

L = O.lib()
OPTS = O.make_opts()


def photon(a, E, Lz, Q):
    p = O.Photon()
    L.ko_photon_init(O.C.byref(p), a, E, Lz, Q)
    return p


def rhs(p, y):
    yy = (O.C.c_double * 5)(*y)
    dy = (O.C.c_double * 5)()
    L.ko_rhs(O.C.byref(p), yy, dy)
    return np.array(dy)


def test_mu_form_equals_reference_theta_form():
    """On the constraint surface the integrated (r, mu) system is the reference's
    derivatives() (kerr.cpp:331-351) under mu = cos(theta), s = -tau, and BL phi:
    dphi_BL = dphi_KS - a vr / Delta (test_kerr.cpp:645-648)."""
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(400):
        a = rng.uniform(0, 0.998)
        E = rng.uniform(0.5, 1.5)
        Lz = rng.uniform(-7, 7)
        Q = rng.uniform(0, 40)
        r = rng.uniform(2.0, 80.0)
        th = rng.uniform(0.05, math.pi - 0.05)
        c = O.Consts(E, Lz, Q, 0.0)
        R = L.ko_radial_potential(a, c, r)
        T = L.ko_polar_potential(a, c, th)
        if R <= 0 or T <= 0:
            continue
        vr = rng.choice([-1, 1]) * math.sqrt(R)
        vth = rng.choice([-1, 1]) * math.sqrt(T)
        d = L.ko_derivatives(a, c, O.State(0.0, r, th, 0.3, vr, vth))
        p = photon(a, E, Lz, Q)
        m, sn = math.cos(th), math.sin(th)
        y = [r, m, 0.3, -vr, sn * vth]  # backward: w_r = -vr, w_mu = -sin(theta) w_theta = sin(theta) vth
        dy = rhs(p, y)
        delta = r * r - 2 * r + a * a
        want = [-vr, sn * vth, -(d.dphi - a * vr / delta), d.dvr, -math.cos(th) * vth * vth - sn * d.dvtheta]
        np.testing.assert_allclose(dy, want, rtol=1e-11, atol=1e-11 * (1 + np.abs(want).max()))
        checked += 1
    assert checked > 100


def test_second_derivatives_are_half_potential_gradients():
    rng = np.random.default_rng(12)
    for _ in range(200):
        a, E, Lz, Q = rng.uniform(0, 0.998), rng.uniform(0.5, 1.5), rng.uniform(-7, 7), rng.uniform(-2, 40)
        p = photon(a, E, Lz, Q)
        r, m = rng.uniform(1.5, 60), rng.uniform(-0.99, 0.99)
        dy = rhs(p, [r, m, 0.0, 0.0, 0.0])
        hr, hm = 1e-3 * r, 1e-4
        Rp = (-L.ko_R(p, r + 2 * hr) + 8 * L.ko_R(p, r + hr) - 8 * L.ko_R(p, r - hr) + L.ko_R(p, r - 2 * hr)) / (12 * hr)
        Mp = (-L.ko_M(p, m + 2 * hm) + 8 * L.ko_M(p, m + hm) - 8 * L.ko_M(p, m - hm) + L.ko_M(p, m - 2 * hm)) / (12 * hm)
        assert dy[3] == pytest.approx(Rp / 2, rel=1e-8, abs=1e-8 * r ** 3)
        assert dy[4] == pytest.approx(Mp / 2, rel=1e-7, abs=1e-8 * (abs(Q) + Lz * Lz + 1))


def test_launches_lie_on_the_constraint_surface():
    rng = np.random.default_rng(13)
    v = O.make_view()
    for _ in range(100):
        al, be = rng.uniform(-12, 12, 2)
        p = O.Photon()
        y0 = (O.C.c_double * 5)()
        L.ko_launch_bardeen(O.C.byref(v), al, be, O.C.byref(p), y0)
        R, M = L.ko_R(p, y0[0]), L.ko_M(p, y0[1])
        assert y0[3] ** 2 == pytest.approx(max(R, 0), rel=1e-14, abs=1e-9)
        assert y0[4] ** 2 == pytest.approx(max(M, 0), rel=1e-14, abs=1e-12)
        assert y0[3] <= 0.0 and (y0[4] > 0) == (be > 0)
    vp = O.make_view(camera_kind=O.PINHOLE, spin=0.94, r_cam=32.0, theta_cam=1.48, yaw=0.3, pitch=-0.1, roll=0.2,
                     width=40, height=30)
    for col in range(0, 40, 7):
        for row in range(0, 30, 7):
            p = O.Photon()
            y0 = (O.C.c_double * 5)()
            L.ko_launch_pixel(O.C.byref(vp), col, row, O.C.byref(p), y0)
            R, M = L.ko_R(p, y0[0]), L.ko_M(p, y0[1])
            P = p.E * (y0[0] ** 2 + p.a2) - p.aL
            assert abs(y0[3] ** 2 - R) < 1e-12 * P * P
            assert abs(y0[4] ** 2 - M) < 1e-10 * (abs(p.Q) + p.L ** 2 + 1)


def test_constraints_hold_at_termination():
    """w_r^2 = R and w_mu^2 = M where rays end (captured or on the disk)."""
    v = O.make_view(width=32, height=32)
    n_checked = 0
    for col in range(0, 32, 3):
        for row in range(0, 32, 3):
            p = O.Photon()
            y0 = (O.C.c_double * 5)()
            L.ko_launch_pixel(O.C.byref(v), col, row, O.C.byref(p), y0)
            res = O.Result()
            L.ko_trace_photon(O.C.byref(v), O.C.byref(OPTS), O.C.byref(p), y0, O.C.byref(res))
            if res.cls not in (O.CAPTURED, O.DISK):
                continue
            y = list(res.y_end)
            P = p.E * (y[0] ** 2 + p.a2) - p.aL
            delta = y[0] ** 2 - 2 * y[0] + p.a2
            assert abs(y[3] ** 2 - L.ko_R(p, y[0])) < 1e-8 * (P * P + abs(delta) * p.K)
            assert abs(y[4] ** 2 - L.ko_M(p, y[1])) < 1e-8 * (abs(p.Q) + p.L ** 2 + p.A2E2)
            n_checked += 1
    assert n_checked > 50


@pytest.mark.parametrize("b", [6.0, 7.5, 10.0, 20.0])
def test_schwarzschild_deflection_exact(b):
    """a = 0, edge-on camera, beta = 0: phi at infinity equals the exact quadrature (1e-9 rad)."""
    r_cam = 50.0
    v = O.make_view(spin=0.0, r_cam=r_cam, theta_cam=math.pi / 2, disk_enabled=0)
    out = O.trace_params(v, OPTS, [[b, 0.0]])
    assert out["cls"][0] == O.ESCAPED
    inward, outward, _ = Qd.schwarzschild_sweep(b, r_cam)
    want = (inward + outward) % (2 * math.pi)
    assert abs(out["phi"][0] - want) < 1e-9, out["phi"][0] - want
    assert abs(out["theta"][0] - math.pi / 2) < 1e-12


@pytest.mark.parametrize("b", [5.4, 5.7, 6.2])
def test_disk_hit_radius_and_redshift_exact(b):
    """a = 0, camera on the equator, alpha = 0 (L = 0: the ray crosses the pole, phi + pi),
    beta = b: it meets the far side of the equatorial disk where the swept angle is pi.
    r_hit against the quadrature; g = sqrt(1 - 3/r) for L = 0 Keplerian gas."""
    r_cam = 30.0
    r_want = Qd.schwarzschild_pole_disk_hit(b, r_cam)
    # the crossing lies just past periapsis, inside r_ISCO = 6: the disk reaches in to 3.05 M
    v = O.make_view(spin=0.0, r_cam=r_cam, theta_cam=math.pi / 2, disk_enabled=1, disk_inner=3.05, disk_outer=40.0)
    out = O.trace_params(v, OPTS, [[0.0, b]])
    assert out["cls"][0] == O.DISK, out
    assert out["r_hit"][0] == pytest.approx(r_want, rel=1e-9)
    assert out["g"][0] == pytest.approx(math.sqrt(1.0 - 3.0 / out["r_hit"][0]), rel=1e-14)


@pytest.mark.parametrize("alpha,beta", [(0.0, -3.0), (3.0, -4.0), (-5.0, -2.0), (2.0, -7.0), (6.0, -1.5),
                                        (-8.0, -3.0)])
def test_kerr_disk_hit_radius_exact(alpha, beta):
    """a = 0.9, Bardeen camera at r = 50 M, i = 60 deg, beta < 0: the ray runs from theta = i
    down to the equator with neither r nor theta turning, so at the crossing the Mino
    times agree: int_i^(pi/2) dtheta / sqrt(Theta) = int_(r_hit)^(r_cam) dr / sqrt(R).
    r_hit from that equation (quadrature + root) to 1e-9 relative; dropping the spin term
    of Theta moves it by ~3e-3.  g is the Keplerian prograde closed form at r_hit."""
    a, r_cam, inc = 0.9, 50.0, math.radians(60.0)
    r_want = Qd.kerr_disk_hit(a, r_cam, inc, alpha, beta)
    v = O.make_view(spin=a, r_cam=r_cam, theta_cam=inc, disk_enabled=1, disk_inner=1.5, disk_outer=100.0)
    out = O.trace_params(v, OPTS, [[alpha, beta]])
    assert out["cls"][0] == O.DISK, out
    r = out["r_hit"][0]
    assert r == pytest.approx(r_want, rel=1e-9)
    assert out["g"][0] == pytest.approx(Qd.keplerian_g(a, r, -alpha * math.sin(inc)), rel=1e-12)


def test_equatorial_critical_impact_parameters():
    """testShadow (test_kerr.cpp:222-262): b = b_c (1 -+ 1e-3) falls in / escapes, for a = 0
    (b_c = 3 sqrt 3) and a = 0.9 (xi(r_ph) = 2.844 prograde, -6.834 retrograde)."""
    def xi(a, r):
        return -(r ** 3 - 3 * r * r + a * a * r + a * a) / (a * (r - 1.0))

    cases = [(0.0, 3 * math.sqrt(3)), (0.0, -3 * math.sqrt(3))]
    a = 0.9
    bp, br = xi(a, L.ko_photon_orbit_radius(a, 1)), xi(a, L.ko_photon_orbit_radius(a, 0))
    assert bp == pytest.approx(2.844, abs=2e-3) and br == pytest.approx(-6.834, abs=2e-3)
    cases += [(a, bp), (a, br)]
    for a_, bc in cases:
        v = O.make_view(spin=a_, r_cam=1000.0, theta_cam=math.pi / 2, disk_enabled=0)
        # L = -alpha sin i = b  ->  alpha = -b
        out = O.trace_params(v, OPTS, [[-bc * 0.999, 0.0], [-bc * 1.001, 0.0]])
        assert out["cls"][0] == O.CAPTURED and out["cls"][1] == O.ESCAPED, (a_, bc, out["cls"])


def exact_critical_radius(a, inc, psi):
    """Distance from the image origin to the critical curve along psi, solving for the
    spherical photon orbit radius r (the curve's parameter, eht.cpp:52-80) exactly."""
    if a < 1e-9:
        return math.sqrt(27.0)
    sI, cI = math.sin(inc), math.cos(inc)

    def pt(r):
        xi = (r * r * (3 - r) - a * a * (r + 1)) / (a * (r - 1))
        eta = r ** 3 * (4 * a * a - r * (r - 3) ** 2) / (a * a * (r - 1) ** 2)
        b2 = eta + a * a * cI * cI - xi * xi * cI * cI / (sI * sI)
        return -xi / sI, math.sqrt(max(b2, 0.0))

    r1, r2 = L.ko_photon_orbit_radius(a, 1), L.ko_photon_orbit_radius(a, 0)
    up = psi <= math.pi
    target = psi if up else 2 * math.pi - psi
    rs = np.linspace(r1, r2, 4001)[1:-1]
    ang = np.array([math.atan2(*pt(r)[::-1]) for r in rs])
    ok = np.array([pt(r)[1] > 0 for r in rs])
    j = int(np.argmin(np.where(ok, np.abs(ang - target), np.inf)))
    f = lambda r: math.atan2(*pt(r)[::-1]) - target
    lo, hi = rs[max(j - 1, 0)], rs[min(j + 1, len(rs) - 1)]
    r_star = optimize.brentq(f, lo, hi, xtol=1e-15)
    return math.hypot(*pt(r_star))


@pytest.mark.parametrize("a,deg", [(0.9, 17.0), (0.9, 60.0), (0.0, 40.0)])
def test_shadow_edge_follows_critical_curve(a, deg):
    """testEht (test_kerr.cpp:2281-2304) tightened: the class boundary along 8 directions,
    bisected 40 times, lies on the exact critical curve to 1e-9 (the reference: 2e-3)."""
    inc = math.radians(deg)
    v = O.make_view(spin=a, r_cam=1000.0, theta_cam=inc, disk_enabled=0)
    worst = 0.0
    for k in range(8):
        psi = 2 * math.pi * (k + 0.25) / 8
        lo, hi = 1.0, 12.0
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            out = O.trace_params(v, OPTS, [[mid * math.cos(psi), mid * math.sin(psi)]])
            if out["cls"][0] == O.CAPTURED:
                lo = mid
            else:
                hi = mid
        worst = max(worst, abs(0.5 * (lo + hi) / exact_critical_radius(a, inc, psi) - 1.0))
    assert worst < 1e-9, worst


def test_near_axis_rays_leave_together():
    """testNearAxisPhotons (test_kerr.cpp:283-325): rays passing +-eps either side of the
    spin axis leave together; top-down rays vary smoothly from snapped (L = 0, through
    the pole) to just-missing ones."""
    def direction(th, ph):
        return np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])

    v = O.make_view(camera_kind=O.PINHOLE, width=1, height=1, spin=0.94, r_cam=32.0, theta_cam=1.48, disk_enabled=0,
                    escape_radius=500.0)
    for eps in (0.0011, 0.003):
        looks = []
        for sgn in (1, -1):
            lk = np.array([-1.0, -0.53, sgn * eps])
            looks.append(lk / np.linalg.norm(lk))
        out = O.trace_params(v, OPTS, looks)
        assert (out["cls"] == O.ESCAPED).all()
        d0, d1 = direction(out["theta"][0], out["phi"][0]), direction(out["theta"][1], out["phi"][1])
        assert abs(d0[0] - d1[0]) < 2e-3 and abs(d0[1] - d1[1]) < 1e-2 and abs(d0[2] - d1[2]) < 2e-3
    vt = O.make_view(camera_kind=O.PINHOLE, width=1, height=1, spin=0.94, r_cam=32.0, theta_cam=0.035,
                     disk_enabled=0, escape_radius=500.0)
    prev = None
    snapped = 0
    for k in range(6):
        eps = 1e-6 * 10 ** k
        lk = np.array([-0.8, -0.6, eps])
        out = O.trace_params(vt, OPTS, [lk / np.linalg.norm(lk)])
        snapped += out["consts"][0][1] == 0.0
        assert out["cls"][0] == O.ESCAPED
        d = direction(out["theta"][0], out["phi"][0])
        if prev is not None:
            assert np.linalg.norm(d - prev) < 2e-3 + 2 * eps
        prev = d
    assert snapped >= 1, "the smallest offsets must take the axis snap (kerr.h:115)"


def test_self_convergence():
    """rtol 1e-10 against rtol 1e-13 on rays away from the critical curve: the asymptotic
    sky direction agrees to 1e-8 rad, r_hit and g to 1e-8 relative."""
    v = O.make_view(width=24, height=24)
    tight = O.make_opts(rtol=1e-13, atol=1e-15)
    al, be = O.critical_curve(0.9, math.radians(80), 200)
    rays = []
    for row in range(24):
        for col in range(24):
            a_ = -12 + (col + 0.5)
            b_ = 12 - (row + 0.5)
            rc = O.critical_radius(al, be, math.atan2(b_, a_))
            if abs(math.hypot(a_, b_) / rc - 1) > 0.05:
                rays.append([a_, b_])
    rays = np.array(rays)
    o1 = O.trace_params(v, OPTS, rays)
    o2 = O.trace_params(v, tight, rays)
    assert np.array_equal(o1["cls"], o2["cls"])
    esc = o1["cls"] == O.ESCAPED
    dth = np.abs(o1["theta"][esc] - o2["theta"][esc])
    dph = np.abs(np.angle(np.exp(1j * (o1["phi"][esc] - o2["phi"][esc]))))
    assert dth.max() < 1e-8 and dph.max() < 1e-8, (dth.max(), dph.max())
    disk = o1["cls"] == O.DISK
    np.testing.assert_allclose(o1["r_hit"][disk], o2["r_hit"][disk], rtol=1e-8)
    np.testing.assert_allclose(o1["g"][disk], o2["g"][disk], rtol=1e-8)


def test_equatorial_reflection_symmetry():
    """Camera in the equatorial plane: beta -> -beta mirrors theta -> pi - theta and keeps
    class, phi, r_hit and g."""
    v = O.make_view(spin=0.9, r_cam=30.0, theta_cam=math.pi / 2)
    rng = np.random.default_rng(3)
    ab = rng.uniform(-10, 10, (60, 2))
    o1 = O.trace_params(v, OPTS, ab)
    o2 = O.trace_params(v, OPTS, ab * np.array([1, -1]))
    same = o1["cls"] == o2["cls"]
    assert same.mean() > 0.95
    esc = same & (o1["cls"] == O.ESCAPED)
    np.testing.assert_allclose(o1["theta"][esc], math.pi - o2["theta"][esc], atol=1e-8)
    disk = same & (o1["cls"] == O.DISK)
    np.testing.assert_allclose(o1["r_hit"][disk], o2["r_hit"][disk], rtol=1e-8)


def test_pow_m15_accuracy():
    """The controller's x^(-1/5) (three Newton steps): what matters is that it is the same
    on every IEEE machine; its accuracy (~4e-12) only nudges the next step size."""
    xs = np.logspace(-150, 150, 4001)
    got = np.array([L.ko_pow_m15(x) for x in xs])
    np.testing.assert_allclose(got, xs ** -0.2, rtol=1e-11)
    assert L.ko_pow_m15(float("inf")) == 0.0 and L.ko_pow_m15(0.0) == 1e20


def test_step_budget_exhaustion_is_a_class_not_an_error():
    v = O.make_view()
    res = O.trace_params(v, O.make_opts(max_steps=10), [[3.0, 2.0]])
    assert res["cls"][0] == O.UNFINISHED and res["steps"][0] == 10


@pytest.mark.parametrize("alpha", [-4.0, -8.0, 8.0, 12.0])
def test_kerr_equatorial_deflection_exact(alpha):
    """a = 0.9, edge-on Bardeen camera, beta = 0 (Q = 0, L = -alpha): the sky phi equals the
    exact quadrature of the equatorial orbit equation to 1e-9 rad, for photons on the
    prograde (L = 4, 8) and the retrograde side (L = -8, -12) of the critical impact
    parameters (2.844 and -6.834); frame dragging makes the two sides differ."""
    a, r_cam = 0.9, 50.0
    v = O.make_view(spin=a, r_cam=r_cam, theta_cam=math.pi / 2, disk_enabled=0)
    out = O.trace_params(v, OPTS, [[alpha, 0.0]])
    assert out["cls"][0] == O.ESCAPED
    want = Qd.kerr_equatorial_phi(a, alpha, r_cam)
    d = (out["phi"][0] - want + math.pi) % (2 * math.pi) - math.pi
    assert abs(d) < 1e-9, (out["phi"][0], want)
    assert abs(out["theta"][0] - math.pi / 2) < 1e-12
