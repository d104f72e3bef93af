"""The closed-form solution (kt_closed.hpp, restated in tests/closed_form.py) on the CPU:
its special functions against scipy and their defining identities, its outcomes against
the FP64 oracle (the DP5 integrator at rtol 1e-10) on every ray of small views of the
workloads, and against exact quadratures that need neither method.  The GPU kernel is held
to the oracle directly in tests/test_gpu_closed_form.py."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import special

import closed_form as CF
import quadratures as Qd
from black_simulation_amd.configs import c1_view, c3_views, c5_view
from oracle import oracle as O

FIELDS = ("camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
          "disk_inner", "disk_outer", "disk_prograde", "disk_enabled", "escape_radius", "capture_factor")


def test_carlson_integrals_match_scipy():
    rng = np.random.default_rng(11)
    for _ in range(400):
        x, y, z, p = rng.uniform(0, 4, 4) * 10.0 ** rng.uniform(-3, 1, 4)
        assert abs(CF.rf(x, y, z) / special.elliprf(x, y, z) - 1) < 1e-14
        assert abs(CF.rj(x, y, z, p) / special.elliprj(x, y, z, p) - 1) < 1e-13
        assert abs(CF.rc(x, y) / special.elliprc(x, y) - 1) < 1e-13
    # one zero argument (the complete integrals) and negative parameters (the polar motion)
    for m in (-30.0, -1.0, -1e-3, 0.0, 0.3, 0.9, 0.999999, 1 - 1e-12):
        assert abs(CF.ell_k(m) / special.ellipk(m) - 1) < 1e-14
        for s in (0.1, 0.7, 0.999):
            assert abs(CF.ell_f(s, m) / special.ellipkinc(math.asin(s), m) - 1) < 1e-13


def test_incomplete_pi_against_quadrature():
    from scipy import integrate
    for n, s, m in ((0.3, 0.8, 0.5), (-2.0, 0.6, -3.0), (0.95, 0.99, 0.2), (0.5, 0.9, 0.99)):
        want = integrate.quad(lambda t: 1 / ((1 - n * math.sin(t) ** 2) * math.sqrt(1 - m * math.sin(t) ** 2)),
                              0, math.asin(s), epsabs=1e-15, epsrel=1e-14)[0]
        assert abs(CF.ell_pi(n, s, m) / want - 1) < 1e-12


def test_jacobi_functions():
    rng = np.random.default_rng(12)
    for _ in range(300):
        m = float(rng.choice([rng.uniform(0, 0.99), 1 - 10 ** rng.uniform(-12, -2), 10 ** rng.uniform(-12, -1)]))
        u = float(rng.uniform(-8, 8))
        sn, cn, dn = CF.sncndn(u, m)
        assert abs(sn * sn + cn * cn - 1) < 1e-14 and abs(dn * dn + m * sn * sn - 1) < 1e-14
        # the amplitude: F(am u | m) = u within the first quarter period
        K = special.ellipk(m)
        if 0 < u < K:
            assert abs(special.ellipkinc(math.atan2(sn, cn), m) - u) < 1e-12 * max(1, u)
        if m < 0.99:  # scipy's own ellipj loses accuracy closer to m = 1
            S, Cc, D, _ = special.ellipj(u, m)
            assert max(abs(sn - S), abs(cn - Cc), abs(dn - D)) < 1e-13
    for u in (0.0, 1e-8, 1.0, 10.0, 123.4, -5000.0):
        s, c = CF.qsincos(u)
        assert abs(s - math.sin(u)) < 1e-15 * max(1.0, abs(u)) and abs(c - math.cos(u)) < 1e-15 * max(1.0, abs(u))


def _view_dict(v):
    O.lib().ko_isco_radius.restype = C.c_double
    d = {k: getattr(v, k) for k in ("spin", "theta_cam", "r_cam", "phi_cam", "disk_outer", "capture_factor",
                                    "escape_radius")}
    d["r_in"] = v.disk_inner if v.disk_inner > 0 else O.lib().ko_isco_radius(C.c_double(v.spin), 1)
    return d


def _against_oracle(v, stride=1):
    want = O.render(O.make_view(**{k: getattr(v, k) for k in FIELDS}), O.make_opts(), stride=stride)
    n = len(want["cls"])
    idx = np.arange(n) * stride
    al = -v.half_width + (idx % v.width + 0.5) * 2 * v.half_width / v.width
    be = v.half_height - (idx // v.width + 0.5) * 2 * v.half_height / v.height
    vd = _view_dict(v)
    got = np.array([CF.solve(CF.bardeen_ray(vd, al[i], be[i])) for i in range(n)], dtype=float)
    cls = got[:, 0].astype(int)
    assert (cls >= 0).all(), "declined rays"
    assert np.array_equal(cls, want["cls"])
    disk = cls == CF.DISK
    r_err = np.abs(got[disk, 3] / want["r_hit"][disk] - 1)
    g_err = np.abs(got[disk, 4] / want["g"][disk] - 1)
    esc = cls == CF.ESCAPED
    t1, p1, t2, p2 = got[esc, 1], got[esc, 2], want["theta"][esc], want["phi"][esc]
    u1 = np.stack([np.sin(t1) * np.cos(p1), np.sin(t1) * np.sin(p1), np.cos(t1)])
    u2 = np.stack([np.sin(t2) * np.cos(p2), np.sin(t2) * np.sin(p2), np.cos(t2)])
    defl = 2 * np.arcsin(np.minimum(1, np.linalg.norm(u1 - u2, axis=0) / 2))
    return np.bincount(cls, minlength=3), r_err, g_err, defl


def test_c1_every_ray_matches_the_oracle():
    """a = 0.9, r = 30 M, i = 80 deg, 24x24 M plane at 40x40: every class equal, disk radius
    and g within 1e-8, sky directions within 1e-8 rad (the oracle's DP5 at rtol 1e-10 is
    itself good to ~1e-9)"""
    v = c1_view()
    v.width = v.height = 40
    counts, r_err, g_err, defl = _against_oracle(v)
    assert counts.min() > 50
    assert r_err.max() < 1e-8 and g_err.max() < 1e-8 and defl.max() < 1e-8


@pytest.mark.parametrize("k", [0, 60, 119])
def test_c5_orbit_frames_match_the_oracle(k):
    """frames of the a = 0.998 orbit (r = 40 M, i = 20 -> 85 deg), every 331st pixel of 1080p"""
    counts, r_err, g_err, defl = _against_oracle(c5_view(k), stride=331)
    assert counts[CF.DISK] > 1000 and counts[CF.CAPTURED] > 100
    assert r_err.max() < 1e-8 and g_err.max() < 1e-8
    if len(defl):
        assert defl.max() < 1e-8


@pytest.mark.parametrize("j", [3, 20, 41, 58])
def test_c3_views_match_the_oracle(j):
    """views of the seeded batch (spin, inclination and camera radius drawn) at 20x20"""
    v = c3_views()[j]
    v.width = v.height = 20
    counts, r_err, g_err, defl = _against_oracle(v)
    assert r_err.max(initial=0) < 1e-8 and g_err.max(initial=0) < 1e-8 and defl.max(initial=0) < 1e-8


def test_close_cameras_match_the_oracle():
    """cameras inside the photon region (r = 1.8-3.2 M), where rays start between the roots r2
    and r3 of R as well as beyond r4: every ray with a valid launch (R, M > 0 at the camera)
    answered, classes equal, disk radius and g within 1e-8"""
    n_region2 = 0
    for a, r, inc, hw in ((0.9, 2.6, 1.2, 7.0), (0.998, 1.8, 1.4, 6.0), (0.9, 2.2, 1.5, 5.0)):
        v = O.make_view(width=24, height=24, spin=a, r_cam=r, theta_cam=inc, half_width=hw, half_height=hw)
        want = O.render(v, O.make_opts())
        vd = _view_dict(v)
        for q in range(24 * 24):
            row, col = divmod(q, 24)
            al, be = -hw + (col + 0.5) * 2 * hw / 24, hw - (row + 0.5) * 2 * hw / 24
            ray = CF.bardeen_ray(vd, al, be)
            lam, eta = ray["lam"], ray["eta"]
            P = r * r + a * a - a * lam
            if P * P - (r * r - 2 * r + a * a) * (eta + (lam - a) ** 2) <= 0:
                continue  # no photon leaves the camera with this (alpha, beta)
            A, B = a * a - eta - lam * lam, 2 * (eta + (lam - a) ** 2)
            xi = CF.resolvent_root(A, B, -a * a * eta)
            z = math.sqrt(xi / 2)
            D34 = -0.5 * A - z * z - B / (4 * z)
            n_region2 += D34 >= 0 and r < z + math.sqrt(D34)
            cls, _, _, r_hit, g = CF.solve(ray)
            assert cls == want["cls"][q], (a, r, q)
            if cls == CF.DISK:
                assert abs(r_hit / want["r_hit"][q] - 1) < 1e-8 and abs(g / want["g"][q] - 1) < 1e-8
    assert n_region2 > 30


def test_exact_quadratures():
    """answers fixed without either method: the equatorial sky angle of Kerr (a = 0.9, prograde
    and retrograde edge-on rays of a camera at r = 40 M) and a disk radius by equal Mino times"""
    view = {"spin": 0.9, "theta_cam": math.pi / 2, "r_cam": 40.0, "phi_cam": 0.0, "disk_outer": 20.0,
            "capture_factor": 1.01, "escape_radius": 1000.0, "r_in": 2.320883042}
    for alpha in (-8.0, 8.0, 12.0):
        ray = CF.bardeen_ray(view, alpha, 0.0)
        ray["disk"] = 0  # an edge-on ray runs in the disk's plane
        ray["mu_o"] = 0.0
        ray["eta"] = 0.0
        cls, th, ph, _, _ = CF.solve(ray)
        assert cls == CF.DECLINED  # eta = 0: the polar forms need eta > 0
    # a ray just off the plane: its sky longitude tends to the equatorial one
    for alpha in (-8.0, 8.0):
        ray = CF.bardeen_ray(dict(view, theta_cam=math.pi / 2 - 1e-7), alpha, 1e-7)
        ray["disk"] = 0
        cls, th, ph, _, _ = CF.solve(ray)
        want = Qd.kerr_equatorial_phi(0.9, alpha, 40.0)
        assert cls == CF.ESCAPED and abs((ph - want + math.pi) % (2 * math.pi) - math.pi) < 1e-6
    a, r_cam, inc = 0.9, 40.0, math.radians(60.0)
    for alpha, beta in ((3.0, -9.0), (-4.0, -10.0)):
        view = {"spin": a, "theta_cam": inc, "r_cam": r_cam, "phi_cam": 0.0, "disk_outer": 100.0,
                "capture_factor": 1.01, "escape_radius": 1000.0, "r_in": 1.0}
        cls, _, _, r_hit, g = CF.solve(CF.bardeen_ray(view, alpha, beta))
        want = Qd.kerr_disk_hit(a, r_cam, inc, alpha, beta)
        assert cls == CF.DISK and abs(r_hit / want - 1) < 1e-10
        assert abs(g / Qd.keplerian_g(a, r_hit, -alpha * math.sin(inc)) - 1) < 1e-12


def test_operation_count_of_the_orbit_frames():
    """the per-ray FP64 operation count quoted for the kernel's roofline (bench.py
    CLOSED_FORM_OPS_PER_RAY) is the mean over rays of the orbit frames"""
    import bench

    ops = {"fp64": [], "fp32": []}
    for k in (0, 40, 80, 119):
        v = c5_view(k)
        vd = _view_dict(v)
        idx = np.arange(0, v.rays, 4099)
        for q in idx:
            al = -v.half_width + (q % v.width + 0.5) * 2 * v.half_width / v.width
            be = v.half_height - (q // v.width + 0.5) * 2 * v.half_height / v.height
            ops["fp64"].append(CF.count_ops(CF.bardeen_ray(vd, al, be, count=True)))
            ops["fp32"].append(CF.count_ops(CF.bardeen_ray(vd, al, be, count=True), fp32=True))
    for prec, o in ops.items():
        mean = float(np.mean(o))
        print(f"closed form ({prec}): {mean:.1f} operations per ray (min {min(o)}, max {max(o)}, {len(o)} rays)")
        assert abs(bench.CLOSED_FORM_OPS_PER_RAY[prec] / mean - 1) < 0.05


def test_fp32_pass_leaves_few_rays_unsure_and_agrees_with_fp64():
    """the FP32 pass's tolerances (R_F to 1e-1, the mean to 1e-4, a 1e-5 band at the disk's
    edges): on the orbit frames it leaves about one ray in a thousand to the FP64 pass (the
    escapes, the disk's edges), and classes every other ray as the FP64 form does"""
    unsure = n = 0
    for k in (0, 60, 119):
        v = c5_view(k)
        vd = _view_dict(v)
        for q in range(0, v.rays, 2999):
            al = -v.half_width + (q % v.width + 0.5) * 2 * v.half_width / v.width
            be = v.half_height - (q // v.width + 0.5) * 2 * v.half_height / v.height
            ray = CF.bardeen_ray(vd, al, be)
            c32, c64 = CF.solve(ray, fp32=True)[0], CF.solve(ray)[0]
            n += 1
            if c32 == CF.UNCERTAIN:
                unsure += 1
            else:
                assert c32 == c64
    assert unsure / n < 0.005
