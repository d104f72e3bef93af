"""The oracle's EHT radiative transfer (ko_bessel_k, ko_synchrotron_emissivity, ko_planck_nu,
ko_plasma_at, ko_flow_velocity, ko_eht_trace, ko_eht_flux) pinned three ways:

  - golden vectors written by the reference's own src/eht.cpp (tests/golden/eht_vectors.npz,
    from tests/golden/make_eht_golden.py): the closed-form pieces to 1e-12, the traced
    intensities loosely (the reference steps with RK4 and its own step rule, so a pixel's
    transfer sum differs at the 1e-3 level; captures must agree exactly);
  - the reference's own known answers (tests/test_kerr.cpp:2306-2320): K2 at 1, 0.5, 10 and
    its hot-electron limit, Planck's Rayleigh-Jeans limit, the emissivity's linearity in n_e
    and its cut-off;
  - what the physics fixes independently of the formulas: scipy's K_nu, the gas's u.u = -1,
    Keplerian rotation outside the ISCO, and inside it the ISCO's conserved energy and
    angular momentum along the plunge (closed forms of Bardeen, Press and Teukolsky 1972).
"""
import math
import os

import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "eht_vectors.npz"))
KB, HP, CL = 1.380649e-16, 6.62607015e-27, 2.99792458e10


def bessel(which, x):
    return np.array([O.lib().ko_bessel_k(which, float(v)) for v in np.atleast_1d(x)])


def test_bessel_matches_the_reference_and_scipy():
    x = G["bessel_x"]
    for w in range(3):
        np.testing.assert_allclose(bessel(w, x), G["bessel_k"][w], rtol=1e-13)
        # Abramowitz and Stegun's polynomials are good to a few 1e-7
        np.testing.assert_allclose(bessel(w, x), sp.kv(w, x), rtol=5e-7)


def test_bessel_known_answers():
    # tests/test_kerr.cpp:2307-2310
    assert bessel(2, 1.0)[0] / 1.6248388986351774 == pytest.approx(1.0, abs=1e-5)
    assert bessel(2, 0.5)[0] / 7.5501835512408695 == pytest.approx(1.0, abs=1e-5)
    assert bessel(2, 10.0)[0] / 2.1509817006932767e-05 == pytest.approx(1.0, abs=1e-5)
    assert bessel(2, 1.0 / 50.0)[0] / (2.0 * 50.0 ** 2) == pytest.approx(1.0, abs=1e-3)


def emissivity(row):
    return O.lib().ko_synchrotron_emissivity(*[float(v) for v in row])


def test_emissivity_matches_the_reference():
    got = np.array([emissivity(r) for r in G["emis_in"]])
    want = G["emis"]
    assert np.array_equal(got == 0.0, want == 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0.0)
    assert np.all(got[:10] == 0.0)  # thetaE below 0.05


def test_emissivity_linear_in_density_and_cut_off_above_nu_s():
    # tests/test_kerr.cpp:2313-2320
    j1 = emissivity([230e9, 1e5, 10.0, 10.0, 0.8])
    j2 = emissivity([230e9, 2e5, 10.0, 10.0, 0.8])
    far = emissivity([230e13, 1e5, 10.0, 10.0, 0.8])
    assert abs(j2 / j1 - 2.0) < 1e-12
    assert far < 1e-10 * j1
    for bad in ([230e9, 0.0, 10.0, 10.0, 0.8], [230e9, 1e5, 0.04, 10.0, 0.8], [230e9, 1e5, 10.0, 0.0, 0.8],
                [230e9, 1e5, 10.0, 10.0, 0.0]):
        assert emissivity(bad) == 0.0


def test_emissivity_against_its_defining_formula_with_scipy():
    # Leung, Gammie and Noble (2011) eq. 72 written out with scipy's exact K2
    rng = np.random.default_rng(3)
    for _ in range(50):
        nu, ne, th, B, s = 10 ** rng.uniform(10, 12), 10 ** rng.uniform(3, 7), 10 ** rng.uniform(-0.5, 1.5), \
            10 ** rng.uniform(-1, 2), rng.uniform(0.2, 1.0)
        nu_c = 4.80320425e-10 * B / (2 * math.pi * 9.1093837015e-28 * CL)
        nu_s = 2.0 / 9.0 * nu_c * th * th * s
        X = nu / nu_s
        want = ne * math.sqrt(2) * math.pi * 4.80320425e-10 ** 2 * nu_s / (3 * sp.kv(2, 1 / th) * CL) \
            * (X ** 0.5 + 2 ** (11 / 12) * X ** (1 / 6)) ** 2 * math.exp(-X ** (1 / 3))
        if want > 1e-300:
            assert emissivity([nu, ne, th, B, s]) == pytest.approx(want, rel=2e-6)


def test_planck_matches_the_reference_and_its_limits():
    got = np.array([O.lib().ko_planck_nu(*[float(v) for v in r]) for r in G["planck_in"]])
    np.testing.assert_allclose(got, G["planck"], rtol=1e-13)
    # Rayleigh-Jeans (tests/test_kerr.cpp:2311): brightness temperature c^2 I / (2 k nu^2) = T
    I = O.lib().ko_planck_nu(230e9, 1e10)
    assert CL ** 2 * I / (2 * KB * 230e9 ** 2) / 1e10 == pytest.approx(1.0, abs=1e-3)
    # Wien: h nu >> k T
    nu, T = 1e16, 1e4  # h nu / k T = 48
    wien = 2 * HP * nu ** 3 / CL ** 2 * math.exp(-HP * nu / (KB * T))
    assert O.lib().ko_planck_nu(nu, T) == pytest.approx(wien, rel=1e-3)


def flow_of(tag):
    return O.flow(*G[f"flow_{tag}"].tolist())


@pytest.mark.parametrize("tag", ["default", "thin"])
def test_plasma_and_velocity_match_the_reference(tag):
    f = flow_of(tag)
    pl = np.array([O.plasma_at(f, r, th, a) for a, r, th in G["flow_geo"]])
    u = np.array([O.flow_velocity(f, a, r, th) for a, r, th in G["flow_geo"]])
    # the oracle works in mu = cos(theta) and takes rho from sqrt(1 - mu^2): off the plane,
    # where the Gaussian's exponent reaches -100, its few-ulp difference is multiplied by it
    np.testing.assert_allclose(pl, G[f"plasma_{tag}"], rtol=2e-10, atol=0.0)
    np.testing.assert_allclose(u, G[f"velocity_{tag}"], rtol=1e-11, atol=1e-14)


def metric(a, r, th):
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    sig, dl = r * r + a * a * c2, r * r - 2 * r + a * a
    A = (r * r + a * a) ** 2 - a * a * dl * s2
    return -(1 - 2 * r / sig), -2 * a * r * s2 / sig, A * s2 / sig, sig / dl


def isco_el(a):
    # prograde ISCO radius, energy and angular momentum (Bardeen, Press and Teukolsky 1972)
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = math.sqrt(3 * a * a + z1 * z1)
    r = 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))
    E = math.sqrt(1 - 2 / (3 * r))
    L = 2 / (3 * math.sqrt(3)) * (1 + 2 * math.sqrt(3 * r - 2))
    if a > 0:  # general a: the circular-orbit formulas at r_isco
        sr = math.sqrt(r)
        den = r ** 0.75 * math.sqrt(r ** 1.5 - 3 * sr + 2 * a)
        E = (r ** 1.5 - 2 * sr + a) / den
        L = (r * r - 2 * a * sr + a * a) / den
    return r, E, L


def test_gas_velocity_is_a_timelike_unit_vector_with_the_right_rotation():
    f = flow_of("default")
    inside = 0
    for a, r, th in G["flow_geo"]:
        u = O.flow_velocity(f, a, r, th)
        gtt, gtp, gpp, grr = metric(a, r, th)
        norm = gtt * u[0] ** 2 + 2 * gtp * u[0] * u[3] + gpp * u[3] ** 2 + grr * u[1] ** 2
        assert norm == pytest.approx(-1.0, abs=1e-10)
        r_isco, E, L = isco_el(a)
        if r >= r_isco:
            assert u[1] == 0.0
            assert u[3] / u[0] == pytest.approx(0.9 / (r ** 1.5 + a), rel=1e-12)
        elif u[1] != 0.0:
            inside += 1
            assert u[1] < 0.0  # falling in
    assert inside > 10
    # in the plane, inside the ISCO, the gas follows the ISCO's plunging geodesic: its
    # energy and angular momentum are the ISCO's, and u^r is the geodesic's
    for a in (0.0, 0.5, 0.9, 0.998):
        r_isco, E, L = isco_el(a)
        r_h = 1 + math.sqrt(1 - a * a)
        for r in np.linspace(r_h * 1.05, r_isco * 0.999, 7):
            u = O.flow_velocity(f, a, r, math.pi / 2)
            gtt, gtp, gpp, grr = metric(a, r, math.pi / 2)
            assert -(gtt * u[0] + gtp * u[3]) == pytest.approx(E, rel=1e-9)
            assert gtp * u[0] + gpp * u[3] == pytest.approx(L, rel=1e-9)
            # (r^2 u^r)^2 = R(r) of the ISCO's E, L with Q = 0 (Carter), timelike
            P = E * (r * r + a * a) - a * L
            R = P * P - (r * r - 2 * r + a * a) * (r * r + (L - a * E) ** 2)
            assert (r * r * u[1]) ** 2 == pytest.approx(max(R, 0.0), rel=1e-8, abs=1e-12)
    # a = 0 at r = 6 exactly: the Schwarzschild ISCO's (E, L) = (sqrt(8/9), sqrt(12))
    r, E, L = isco_el(0.0)
    assert (r, E, L) == pytest.approx((6.0, math.sqrt(8 / 9), math.sqrt(12)), rel=1e-12)


def test_plasma_profile_laws():
    f = O.flow(1e5, 4e11, 10.0, 0.9, 1.0, 0.8165)
    a = 0.5
    r_isco = isco_el(a)[0]
    for r in (8.0, 15.0, 40.0):
        ne, Te, B = O.plasma_at(f, r, math.pi / 2, a)
        assert ne == pytest.approx(1e5 * r ** -1.1, rel=1e-13)
        assert Te == pytest.approx(4e11 * r ** -0.84, rel=1e-13)
        assert B == pytest.approx(math.sqrt(8 * math.pi * ne * 1.67262192369e-24 * CL ** 2 / (60 * r)), rel=1e-13)
        # Gaussian in z = r cos(theta) of width h rho off the plane
        th = 1.2
        z, rho = r * math.cos(th), r * math.sin(th)
        assert O.plasma_at(f, r, th, a)[0] == pytest.approx(ne * math.exp(-z * z / (2 * rho * rho)), rel=1e-12)
    # inside the ISCO the density thins as the gas speeds up (continuity): below r^-1.1
    r = 0.5 * (r_isco + 1.0 + math.sqrt(1 - a * a))
    assert O.plasma_at(f, r, math.pi / 2, a)[0] < 1e5 * r ** -1.1


def test_traced_intensities_follow_the_reference():
    trace = G["trace"]
    for view, flow, rays in zip(G["trace_views"], G["trace_flows"], trace):
        a, inc, hw, nu, mass = view
        v = O.eht_view(a, inc, hw, 8, nu, mass)
        f = O.flow(*flow.tolist())
        o = O.make_opts()
        got = np.array([[O.eht_trace(v, f, o, al, b).intensity, O.eht_trace(v, f, o, al, b).captured]
                        for al, b in rays[:, :2]])
        assert np.array_equal(got[:, 1], rays[:, 3])
        want = rays[:, 2]
        lit = want > 1e-3 * want.max()
        rel = np.abs(got[lit, 0] / want[lit] - 1.0)
        print(f"a {a}: {lit.sum()} lit rays, rel err median {np.median(rel):.2e} max {rel.max():.2e}; "
              f"total ratio {got[:, 0].sum() / want.sum():.4f}")
        assert np.median(rel) < 5e-3 and rel.max() < 3e-2
        assert got[:, 0].sum() / want.sum() == pytest.approx(1.0, abs=1.5e-2)
        assert np.all(got[rays[:, 3] == 1, 0] >= 0.0)


def test_trace_transfer_converges_with_the_step_tolerance():
    # The transfer sum is the reference's rectangle rule over the steps, so it converges
    # at first order in the step: tightening rtol 100 times moves it by well under 1%.
    v = O.eht_view(0.9, math.radians(60.0), 12.0, 8, 230e9, 6.5e9)
    f = O.flow(1e5, 4e11, 10.0, 0.9, 1.0, 0.8165)
    for al, b in [(7.0, 3.0), (-6.0, 4.0), (1.0, -7.0), (2.0, 1.0)]:
        r1 = O.eht_trace(v, f, O.make_opts(1e-10, 1e-12), al, b)
        r2 = O.eht_trace(v, f, O.make_opts(1e-12, 1e-14), al, b)
        assert r1.captured == r2.captured
        assert r1.intensity > 0.0
        assert r2.intensity == pytest.approx(r1.intensity, rel=1e-2)
        assert r2.steps > r1.steps


def test_small_image_flux_doppler_and_shadow():
    """tests/test_kerr.cpp:2322-2356 at 24 x 24: the approaching side (alpha < 0) is the
    brighter, the shadow's area is the critical curve's, and ko_eht_flux is the sum of the
    intensities times the pixel's solid angle in janskys."""
    v = O.eht_view(0.9, math.radians(60.0), 16.0, 24, 230e9, 6.5e9)
    f = O.flow(1e5, 4e11, 10.0, 0.9, 1.0, 0.8165)
    inten, shadow = O.eht_render(v, f, O.make_opts())
    left, right = inten[:, :12].sum(), inten[:, 12:].sum()
    assert left > 1.5 * right
    px = 2 * 16.0 / 24
    al, be = O.critical_curve(0.9, math.radians(60.0), 400)
    area = 0.5 * abs(np.sum(al * np.roll(be, -1) - np.roll(al, -1) * be))
    assert shadow.sum() * px * px / area == pytest.approx(1.0, abs=0.08)
    theta_g = 3.81899315
    uas = math.pi / 180 / 3600 / 1e6
    want = inten.sum() * (px * theta_g * uas) ** 2 / 1e-23
    assert O.eht_flux(inten, 16.0, theta_g) == pytest.approx(want, rel=1e-12)


def test_calibrated_density_matches_the_reference():
    """eht::calibrate (eht.cpp:432-457) on M87*'s 0.5 Jy for a = 0.9 seen from 163 degrees:
    the secant lands within 1% of the flux, so the two n0 agree to about that."""
    ci = G["calibrate_in"]
    v = O.eht_view(ci[0], ci[1], ci[2], 48, ci[3], ci[4])
    f = O.flow(*G["calibrate_flow"].tolist())
    n0 = O.eht_calibrate(v, f, O.make_opts(), ci[5], ci[6])
    assert n0 == pytest.approx(float(G["calibrate_n0"]), rel=1e-2)
    f.n0 = n0
    inten, _ = O.eht_render(v, f, O.make_opts())
    assert O.eht_flux(inten, ci[2], ci[6]) == pytest.approx(ci[5], rel=1e-2)
