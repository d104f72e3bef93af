"""Naked singularities (1 < a <= 1.5, src/kerr.cpp:117 and its naked branches): the oracle's
radii, Kerr-Schild shift and traced rays against the reference's own outputs
(tests/golden/naked_vectors.npz, written by tests/golden/make_naked_golden.py from the
reference's code compiled in place) and against what the mathematics fixes, restating the
reference's testNakedSingularity (tests/test_kerr.cpp:2456-2507)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "naked_vectors.npz"))
L = O.lib()
for _f, _args in (("ko_isco_radius", [C.c_double, C.c_int]), ("ko_photon_orbit_radius", [C.c_double, C.c_int]),
                  ("ko_horizon_radius", [C.c_double]), ("ko_ks_shift", [C.c_double, C.c_double])):
    getattr(L, _f).restype = C.c_double
    getattr(L, _f).argtypes = _args


def circular_energy(a, r, prograde):
    """E of the circular equatorial orbit at r (Bardeen, Press & Teukolsky 1972)"""
    s, x = (1.0 if prograde else -1.0), math.sqrt(r)
    return (r * x - 2.0 * x + s * a) / (r ** 0.75 * math.sqrt(r * x - 3.0 * x + 2.0 * s * a))


def test_radii_match_the_reference():
    """no horizon; the retrograde photon orbit as the reference's bisection (1e-13); the ISCO
    within 1e-7 of the reference's golden-section least energy, and a better root of
    dE/dr = 0 than it (the reference's search resolves the flat minimum to ~1e-8)"""
    for a, (rh, isco_p, isco_r, ph_p, ph_r) in zip(G["spins"], G["radii"]):
        assert L.ko_horizon_radius(a) == 0.0 == rh
        assert L.ko_photon_orbit_radius(a, 1) == 0.0 == ph_p
        assert L.ko_photon_orbit_radius(a, 0) == pytest.approx(ph_r, rel=1e-13)
        for prograde, want in ((1, isco_p), (0, isco_r)):
            ours = L.ko_isco_radius(a, prograde)
            assert ours == pytest.approx(want, rel=1e-7), (a, prograde)
            h = 1e-5 * ours
            dE = lambda r: (circular_energy(a, r + h, prograde) - circular_energy(a, r - h, prograde)) / (2 * h)
            assert abs(dE(ours)) <= abs(dE(want)) + 1e-10, (a, prograde)


def test_isco_and_photon_orbit_by_their_defining_equations():
    """testNakedSingularity: the retrograde photon orbit solves r (r - 3)^2 = 4 a^2 and there
    is no prograde one; the ISCO is a least-energy circular orbit, inside r = M prograde"""
    for a in (1.0001, 1.05, 1.2, 1.35, 1.5):
        r = L.ko_photon_orbit_radius(a, 0)
        assert r * (r - 3.0) ** 2 == pytest.approx(4 * a * a, rel=1e-12)
        for prograde in (1, 0):
            ri = L.ko_isco_radius(a, prograde)
            e = lambda x: circular_energy(a, x, prograde)
            assert e(ri) < e(0.9 * ri) or not math.isfinite(e(0.9 * ri))
            assert e(ri) < e(1.1 * ri)
            assert (ri < 1.0) == bool(prograde)


def test_kerr_schild_shift():
    """psi has derivative a / Delta on both sides of r = 0 (testNakedSingularity's check of
    kerr::azimuthShift) and vanishes at infinity; the reference's azimuthShift is the same
    function less a pi / (2 b)"""
    for (a, r), ref in zip(G["azshift_in"], G["azshift"]):
        b = math.sqrt(a * a - 1.0)
        assert L.ko_ks_shift(a, r) == pytest.approx(ref - a * math.pi / (2 * b), abs=1e-13)
        h = 1e-5
        d = (L.ko_ks_shift(a, r + h) - L.ko_ks_shift(a, r - h)) / (2 * h)
        assert d == pytest.approx(a / (r * r - 2 * r + a * a), abs=1e-8)
        assert r * r - 2 * r + a * a >= b * b - 1e-12
    assert abs(L.ko_ks_shift(1.2, 1e12)) < 1e-11


def test_captures_match_the_reference():
    """eht::traceRay from r = 1000 M over three Bardeen planes: a ray is captured where it
    passes through the ring's disk to r < 0 (eht.cpp:285).  Only rays with Q < 0 reach r = 0
    (R(0) = -a^2 Q); those within |Q| < 1e-3 of that boundary are left out (the reference's
    RK4 bottoms out a hair above or below r = 0 there)"""
    opts = O.make_opts()
    for (a, inc, hw), flags in zip(G["eht_planes"], G["eht_captured"]):
        n = flags.shape[0]
        v = O.make_view(width=n, height=n, spin=a, r_cam=1000.0, theta_cam=inc, half_width=hw, half_height=hw,
                        disk_enabled=0, capture_factor=1.0 + 1e-4)
        out = O.render(v, opts)
        ours = (out["cls"] == O.CAPTURED).reshape(n, n)
        checked = 0
        for row in range(n):
            be = hw - (row + 0.5) * 2 * hw / n
            for col in range(n):
                al = -hw + (col + 0.5) * 2 * hw / n
                if abs(be * be + (al * al - a * a) * math.cos(inc) ** 2) < 1e-3:
                    continue
                checked += 1
                assert ours[row, col] == bool(flags[row, col]), (a, inc, row, col)
        assert checked > 0.95 * n * n and flags.sum() >= 10


def test_camera_rays_match_the_reference():
    """trajectory::traceRay from a ZAMO camera at r = 30 M around a = 1.2, with the disk from
    the ISCO (inside r = M) to 20 M: classes equal; disk radius and g within the reference's
    RK4 accuracy (2e-3).  The reference ends a ray that sinks below r = 0.01 M as captured
    (trajectory.cpp:377 with r+ = 0); this tracer, as eht::traceRay, where it passes r = 0:
    a ray that turns in between (its least radius below 0.01 M) is the reference's capture
    and goes on here."""
    a, r, th, r_in, r_out, esc = G["traj_camera"]
    v = O.make_view(camera_kind=O.PINHOLE, width=1, height=1, spin=a, r_cam=r, theta_cam=th, phi_cam=0.0,
                    disk_enabled=1, disk_inner=r_in, disk_outer=r_out, escape_radius=esc, capture_factor=1.01)
    out = O.trace_params(v, O.make_opts(), G["traj_looks"], report=True)
    mapping = {0: O.ESCAPED, 1: O.CAPTURED, 2: O.DISK}
    for k, ref in enumerate(G["traj_out"]):
        want = mapping[int(ref[0])]
        if want == O.CAPTURED and out["cls"][k] != O.CAPTURED:
            assert 0.0 < out["r_min"][k] < 0.01 and ref[4] < 0.01, (k, out["cls"][k], ref[:5])
            continue
        assert out["cls"][k] == want, (k, out["cls"][k], ref[:5])
        if want == O.DISK:
            assert out["r_hit"][k] == pytest.approx(ref[1], rel=2e-3)
            assert out["g"][k] == pytest.approx(ref[2], rel=2e-3)
        if want == O.CAPTURED:  # through the ring's disk: the trace ends just past r = 0
            assert -0.5 < out["y_end"][k][0] < 0.0
    assert set(out["cls"]) == {O.ESCAPED, O.CAPTURED, O.DISK}


def test_ray_through_the_rings_disk():
    """testNakedSingularity's ray (E = 1, L = 0, Q = -1.2 at a = 1.2 from r = 8 M, theta = 0.3):
    it reaches r = 0 away from the equator and passes through to r < 0, where the trace ends,
    with the constraints kept to the end"""
    a, th = 1.2, 0.3
    beta = math.sqrt(a * a * math.cos(th) ** 2 - 1.2)  # Q = beta^2 - a^2 cos^2 i with alpha = 0
    v = O.make_view(width=1, height=1, spin=a, r_cam=8.0, theta_cam=th, half_width=1.0, half_height=1.0,
                    disk_enabled=0, capture_factor=1.01)
    out = O.trace_params(v, O.make_opts(), [[0.0, beta]])
    E, Lz, Q = out["consts"][0]
    assert Lz == 0.0 and Q == pytest.approx(-1.2, abs=1e-12)
    assert out["cls"][0] == O.CAPTURED
    r_end, mu_end, _, wr, wm = out["y_end"][0]
    assert -0.5 < r_end < 0.0 and abs(mu_end) > 0.2  # through the disk r = 0, away from the ring
    assert out["steps"][0] < 5000
    P, K = r_end * r_end + a * a - a * Lz, (Lz - a) ** 2 + Q
    R = P * P - (r_end * r_end - 2 * r_end + a * a) * K
    M = Q - (Q + Lz * Lz - a * a) * mu_end ** 2 - a * a * mu_end ** 4
    assert wr * wr == pytest.approx(R, rel=1e-8, abs=1e-10) and wm * wm == pytest.approx(M, rel=1e-8, abs=1e-10)
