"""The oracle's observer cameras (KO_CAMERA_OBSERVER): the pinhole camera carried by a
boosted ZAMO or a free-falling observer, inside the horizon too, integrated with the
ingoing Kerr-Schild phi.  Pinned by

  - golden vectors from the reference (tests/golden/observer_vectors.npz, from
    tests/golden/make_observer_golden.py): kerr::derivatives' Kerr-Schild dphi inside and
    near the horizon, kerr::fourVelocity, and trajectory::traceRay from boosted and falling
    observers' frames, including the far-side rule;
  - what the physics fixes: a ZAMO observer view equals the ZAMO pinhole view up to
    phi_KS - phi_BL at the camera; a boost is special-relativistic aberration and Doppler
    (for boosts along a frame axis, where Gram-Schmidt gives the pure boost); a falling
    camera sees continuously across the horizon; a radial ray seen from inside the
    Schwarzschild horizon comes straight from the camera's own direction.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

G = np.load(__file__.rsplit("/", 1)[0] + "/golden/observer_vectors.npz")
L = O.lib()
OPTS = O.make_opts()


def rhs(p, y):
    yy = (O.C.c_double * 5)(*y)
    dy = (O.C.c_double * 5)()
    L.ko_rhs(O.C.byref(p), yy, dy)
    return np.array(dy)


def test_ks_phi_equals_reference_derivatives_inside_and_near_the_horizon():
    """dphi/ds of the KS integrator is -dphi/dtau of kerr::derivatives (kerr.cpp:331-351)
    for photon states inside the horizon, within 1e-6..1e-2 of it, and outside."""
    for (a, E, Lz, Q, r, th, vr, vth), d in zip(G["ksrhs_in"], G["ksrhs"]):
        p = O.Photon()
        L.ko_photon_init(O.C.byref(p), a, E, Lz, Q)
        p.ks = 1
        m, sn = math.cos(th), math.sin(th)
        dy = rhs(p, [r, m, 0.3, -vr, sn * vth])
        scale = 1.0 + abs(d[3]) + abs(Lz) / sn ** 2 + abs(a)
        assert dy[2] == pytest.approx(-d[3], rel=1e-9, abs=1e-10 * scale), (a, r, th)
        assert dy[0] == -vr and dy[3] == pytest.approx(d[4], rel=1e-11, abs=1e-9 * (1 + abs(d[4])))


def four_velocity(a, E, Lz, Q, r, th, sr, st):
    c = O.Consts(E, Lz, Q, 1.0)
    R = L.ko_radial_potential(a, c, r)
    T = L.ko_polar_potential(a, c, th)
    d = L.ko_derivatives(a, c, O.State(0.0, r, th, 0.0, sr * math.sqrt(max(R, 0.0)), st * math.sqrt(max(T, 0.0))))
    sig = r * r + a * a * math.cos(th) ** 2
    return np.array([d.dt, d.dr, d.dtheta, d.dphi]) / sig


def test_geodesic_four_velocity_matches_reference():
    """kerr::fourVelocity (kerr.cpp:414-418) through the oracle's ko_derivatives, and the
    library's kt_geodesic_four_velocity against the same golden values."""
    from black_simulation_amd import _lib as K

    u_lib = (K.C.c_double * 4)()
    for (a, E, Lz, Q, r, th, sr, st), want in zip(G["fourvel_in"], G["fourvel"]):
        got = four_velocity(a, E, Lz, Q, r, th, sr, st)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
        assert K.lib().kt_geodesic_four_velocity(a, E, Lz, Q, r, th, int(sr), int(st), u_lib) == K.KT_OK
        np.testing.assert_allclose(np.array(u_lib), want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
        # timelike and normalised: g(u, u) = -1 in the Kerr-Schild metric
        g = np.zeros((4, 4))
        L.ko_metric_tensor(a, r, th, g.ctypes.data_as(O.C.POINTER(O.C.c_double)))
        assert want @ g @ want == pytest.approx(-1.0, abs=1e-9)


def observer_view(a, r, th, u=(0, 0, 0, 0), drift=(0, 0, 0), disk=0, rin=0.0, rout=20.0, kind=O.OBSERVER,
                  capture=None):
    rp = L.ko_horizon_radius(a)
    return O.make_view(camera_kind=kind, width=1, height=1, spin=a, r_cam=r, theta_cam=th, phi_cam=0.0,
                       disk_enabled=disk, disk_inner=rin, disk_outer=rout, escape_radius=1000.0,
                       capture_factor=capture or 1.0 + 0.01 / rp, u_obs=u, drift=drift)


def sky(out, k):
    th, ph = out["theta"][k], out["phi"][k]
    return np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])


def test_traces_from_observer_frames_match_reference():
    """trajectory::traceRay (trajectory.cpp:214-396) from boosted ZAMOs and from cameras
    falling from rest at infinity at r = 1.2 M, on either side of the horizon and at 1.6 M
    (a = 0.9), with a general falling camera inside and a Schwarzschild one: the same class
    for all 160 rays (the far-side rule included), the sky direction to 1e-5 rad, the
    photon's E and L/E, and the disk's r and g to the reference's RK4 accuracy."""
    n_dark = 0
    for case, ref in zip(G["trace_cases"], G["trace_out"]):
        a, r, th, u0, u1, u2, u3, lx, ly, lz, disk, rin, rout, esc = case
        out = O.trace_params(observer_view(a, r, th, u=(u0, u1, u2, u3), disk=int(disk), rin=rin, rout=rout), OPTS,
                             [[lx, ly, lz]])
        ours, want = int(out["cls"][0]), int(ref[0])
        assert ours == {0: O.ESCAPED, 1: O.CAPTURED, 2: O.DISK}[want], (case, ours, want)
        E, Lz, Q = out["consts"][0]
        assert E == pytest.approx(ref[8], rel=1e-11, abs=1e-12)
        assert Lz / E == pytest.approx(ref[9], rel=1e-10, abs=1e-10)
        if want == 0:
            assert 1.0 - np.dot(sky(out, 0), ref[5:8]) < 5e-11
        if want == 2:
            assert out["r_hit"][0] == pytest.approx(ref[1], rel=2e-3)
            assert out["g"][0] == pytest.approx(ref[2], rel=2e-3)
        rp = L.ko_horizon_radius(a)
        if r < rp and 2 * rp * E - a * Lz < 0:
            n_dark += 1
            assert ours == O.CAPTURED and out["steps"][0] == 0
    assert n_dark >= 3


LOOKS = np.array([[1, 0, 0], [-0.3, 0.9, 0.2], [0.2, -0.4, 0.8], [-0.6, -0.1, -0.7], [0.5, 0.5, -0.7],
                  [0.05, 0.1, 0.99], [-0.9, 0.3, -0.3], [0.7, -0.7, 0.1]], float)
LOOKS /= np.linalg.norm(LOOKS, axis=1, keepdims=True)


def test_static_observer_is_the_zamo_camera_up_to_the_ks_phi_shift():
    """With no boost the observer view is the ZAMO pinhole view; only phi differs, by
    phi_KS - phi_BL = psi(r_cam) (zero at infinity), since the observer view's phi_cam is
    the Kerr-Schild phi."""
    a, r, th = 0.9, 8.0, 1.3
    isco = L.ko_isco_radius(a, 1)
    zamo = O.trace_params(observer_view(a, r, th, disk=1, rin=isco, kind=O.PINHOLE), OPTS, LOOKS)
    obs = O.trace_params(observer_view(a, r, th, disk=1, rin=isco), OPTS, LOOKS)
    psi = L.ko_ks_shift(a, r)
    assert psi == pytest.approx(0.9 / (2 * math.sqrt(1 - 0.81)) * math.log(
        (r - 1 - math.sqrt(0.19)) / (r - 1 + math.sqrt(0.19))), rel=1e-13)
    assert np.array_equal(zamo["cls"], obs["cls"])
    n_sky = 0
    for k in range(len(LOOKS)):
        if zamo["cls"][k] == O.ESCAPED:
            n_sky += 1
            assert obs["theta"][k] == pytest.approx(zamo["theta"][k], abs=1e-8)
            dphi = (obs["phi"][k] - zamo["phi"][k] + psi + math.pi) % (2 * math.pi) - math.pi
            assert abs(dphi) < 1e-8
        if zamo["cls"][k] == O.DISK:
            assert obs["r_hit"][k] == pytest.approx(zamo["r_hit"][k], rel=1e-8)
            assert obs["g"][k] == pytest.approx(zamo["g"][k], rel=1e-8)
    assert n_sky >= 3


@pytest.mark.parametrize("axis,v", [(0, -0.5), (1, 0.3), (2, 0.6), (2, -0.8)])
def test_boost_is_aberration_and_doppler(axis, v):
    """A ZAMO boosted with speed v along frame axis `axis` (r, theta, phi) sees along n' the
    photon the ZAMO sees along n = (n'_x - v, n'_perp / gamma) / (1 - v n'_x) (aberration;
    n is where the photon comes from), and measures it gamma (1 + v n_x) times bluer, so
    its disk g is the ZAMO's times that factor (Doppler)."""
    a, r, th = 0.9, 10.0, 1.2
    isco = L.ko_isco_radius(a, 1)
    gamma = 1.0 / math.sqrt(1.0 - v * v)
    drift = [0.0, 0.0, 0.0]
    drift[axis] = v
    zl = []
    for n1 in LOOKS:
        x = n1[axis]
        n = n1 / (gamma * (1.0 - v * x))
        n[axis] = (x - v) / (1.0 - v * x)
        zl.append(n / np.linalg.norm(n))
    zl = np.array(zl)
    boosted = O.trace_params(observer_view(a, r, th, drift=drift, disk=1, rin=isco), OPTS, LOOKS)
    zamo = O.trace_params(observer_view(a, r, th, disk=1, rin=isco), OPTS, zl)
    assert np.array_equal(boosted["cls"], zamo["cls"])
    for k in range(len(LOOKS)):
        if zamo["cls"][k] == O.ESCAPED:
            assert 1.0 - np.dot(sky(boosted, k), sky(zamo, k)) < 1e-14
        if zamo["cls"][k] == O.DISK:
            assert boosted["r_hit"][k] == pytest.approx(zamo["r_hit"][k], rel=1e-9)
            assert boosted["g"][k] == pytest.approx(zamo["g"][k] * gamma * (1.0 + v * zl[k][axis]), rel=1e-9)


def test_falling_camera_sees_continuously_across_the_horizon():
    """A camera falling from rest at infinity (E = 1, L = Q = 0) notices nothing at the
    horizon: at r+(1 -+ 1e-6) it sees the same sky (to 1e-4 rad) and the same classes; the
    rays that inside come from the far side are captured just outside too."""
    a, th = 0.9, 1.0
    rp = L.ko_horizon_radius(a)
    rng = np.random.default_rng(3)
    looks = rng.normal(size=(40, 3))
    looks /= np.linalg.norm(looks, axis=1, keepdims=True)
    res = []
    for r in (rp * (1 - 1e-6), rp * (1 + 1e-6)):
        u = four_velocity(a, 1.0, 0.0, 0.0, r, th, -1, 1)
        res.append(O.trace_params(observer_view(a, r, th, u=tuple(u), capture=1.01), OPTS, looks))
    inside, outside = res
    assert np.array_equal(inside["cls"], outside["cls"])
    assert (inside["cls"] == O.CAPTURED).sum() > 3 and (inside["cls"] == O.ESCAPED).sum() > 10
    for k in np.nonzero(inside["cls"] == O.ESCAPED)[0]:
        assert math.acos(min(1.0, np.dot(sky(inside, k), sky(outside, k)))) < 1e-4


@pytest.mark.parametrize("r", [1.2, 1.6, 1.95])
def test_radial_ray_inside_the_schwarzschild_horizon(r):
    """a = 0, a camera falling from rest inside r = 2: the ray arriving along its outward
    radial direction came straight down from the sky point above it (theta_inf = theta,
    phi_inf = phi, exactly); the one arriving from below has E < 0 and is dark."""
    th = 1.1
    u = four_velocity(0.0, 1.0, 0.0, 0.0, r, th, -1, 1)
    out = O.trace_params(observer_view(0.0, r, th, u=tuple(u), capture=1.01), OPTS, [[1, 0, 0], [-1, 0, 0]])
    assert out["cls"][0] == O.ESCAPED
    assert out["theta"][0] == pytest.approx(th, abs=1e-12)
    assert abs((out["phi"][0] + math.pi) % (2 * math.pi) - math.pi) < 1e-12
    E = out["consts"][1][0]
    assert E < 0 and out["cls"][1] == O.CAPTURED and out["steps"][1] == 0


def oracle_frame(a, r, th, u):
    e = np.zeros(16)
    L.ko_observer_frame(a, r, th, O._dp(np.ascontiguousarray(u, np.float64)), O._dp(e))
    return e.reshape(4, 4)


def test_observer_frames_match_reference():
    """ko_observer_frame against kerr::observerFrame (golden, from the reference's own code)
    for boosted ZAMOs and falling observers outside, at and inside the horizon."""
    for (a, r, th, *u), ref in zip(G["frame_in"], G["frame_out"]):
        got = oracle_frame(a, r, th, u).ravel()
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref))), (a, r, th)


def test_an_explicit_frame_is_the_frame_the_observer_gives():
    """A view given its observer's frame outright (as uCamPhoton carries Camera::frame) traces
    exactly as the view given the 4-velocity; the same frame with its spatial axes turned by
    R, looked at along n', sees what the first sees along R^T n'."""
    a, r, th = 0.9, 1.2, 1.1
    u = four_velocity(a, 1.0, 1.5, 2.0, r, th, -1, 1)
    vu = observer_view(a, r, th, u=tuple(u), disk=1, rin=L.ko_isco_radius(a, 1), capture=1.01)
    un = np.zeros(4)
    L.ko_observer_u(C.byref(vu), O._dp(un))
    e = oracle_frame(a, r, th, un)
    vf = observer_view(a, r, th, disk=1, rin=L.ko_isco_radius(a, 1), capture=1.01)
    vf.frame = (C.c_double * 16)(*e.ravel())
    want = O.trace_params(vu, OPTS, LOOKS)
    got = O.trace_params(vf, OPTS, LOOKS)
    for k in ("cls", "steps", "theta", "phi", "r_hit", "g"):
        assert np.array_equal(got[k], want[k], equal_nan=True), k
    c, s = math.cos(0.7), math.sin(0.7)
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]) @ np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])
    e2 = e.copy()
    e2[1:] = R @ e[1:]
    vf.frame = (C.c_double * 16)(*e2.ravel())
    turned = O.trace_params(vf, OPTS, LOOKS)
    plain = O.trace_params(vu, OPTS, LOOKS @ R)  # rows n^T R = (R^T n)^T
    assert np.array_equal(turned["cls"], plain["cls"])
    for k in np.nonzero(plain["cls"] == O.ESCAPED)[0]:
        assert math.acos(min(1.0, np.dot(sky(turned, k), sky(plain, k)))) < 1e-9
