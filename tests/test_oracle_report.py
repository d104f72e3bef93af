"""The oracle's single-ray report (ko_trace_photon_report): the RayReport fields turns, rMin
and delay of trajectory::traceRay (src/trajectory.h:97-113, src/trajectory.cpp:214-396) as
App::inspectPixel prints them (src/app.cpp:776-777).

Pinned by what the geometry fixes at a = 0, where every ray stays in one plane through the
centre, and by the reference's own traceRay compiled in place (tests/golden/report_vectors.npz):

  * a radial ray's Kerr-Schild delay: dt/dr = (r + 2)/(r - 2) outgoing, the reference's own
    check (test_kerr.cpp:511-518), and dt/dr = -1 for ingoing light, exactly;
  * turns = (the angle swept in the ray's plane) / 2 pi = the integral of
    du / sqrt(1/b^2 - u^2 + 2 u^3), u = 1/r, through the periapsis;
  * r_min = the periapsis, the largest root of r^3 - b^2 r + 2 b^2 = 0;
  * delay = the Boyer-Lindquist time integral du / (b u^2 (1 - 2u) sqrt(...)) plus the
    Kerr-Schild shift 2 ln(r - 2) between the ends.
A dropped term in dt/dtau, a Boyer-Lindquist phi where Kerr-Schild is meant, a missed
half-turn or a periapsis taken from the step ends fails one of them.  Against the
reference, which marches RK4 with steps of 1 % and stops one step past each event, the
bar is its own accuracy."""
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import brentq

from oracle import oracle as O

L = O.lib()
G = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "report_vectors.npz"))


def look_view(a, r, th, disk=0, r_in=2.0, r_out=20.0, esc=250.0, cap=None):
    rh = L.ko_horizon_radius(a)
    return O.make_view(camera_kind=O.PINHOLE, width=1, height=1, spin=a, r_cam=r, theta_cam=th, phi_cam=0.0,
                       disk_enabled=disk, disk_inner=r_in, disk_outer=r_out, escape_radius=esc,
                       capture_factor=cap if cap is not None else 1.0 + 0.01 / rh)


def report(v, look):
    out = O.trace_params(v, O.make_opts(), [look], report=True)
    return {k: (v_[0] if np.ndim(v_) == 1 else v_[0]) for k, v_ in out.items()}


def look_for(b, r):
    """equatorial look of a static camera at r (a = 0) whose ray has impact parameter b"""
    sa = b * math.sqrt(1.0 - 2.0 / r) / r
    return [-math.sqrt(1.0 - sa * sa), 0.0, sa]


def G_u(b, u):
    return 1.0 / b ** 2 - u * u + 2.0 * u ** 3


def periapsis_u(b, r):
    return brentq(lambda u: G_u(b, u), 1.0 / r, 1.0 / 3.0, xtol=1e-17)


def through_periapsis(b, ut, r, f):
    """integral of f(u) / sqrt(G(u)) du from 1/r up to the periapsis u_t (u = u_t - w^2)"""
    g = lambda w: 2.0 * w * f(ut - w * w) / math.sqrt(G_u(b, ut - w * w)) if w > 0 else \
        2.0 * f(ut) / math.sqrt(2.0 * ut - 6.0 * ut * ut)  # noqa: E731
    return quad(g, 0.0, math.sqrt(ut - 1.0 / r), epsabs=1e-14, epsrel=1e-13, limit=400)[0]


def test_radial_delays_are_the_closed_forms():
    """a = 0, camera at r = 30: looking at the hole the ray falls to r_cap with delay
    T(30) - T(r_cap), T(r) = r + 4 ln(r - 2) (the reference's tortoise check,
    test_kerr.cpp:511-518); looking away it climbs to r_esc with delay r_esc - 30 (ingoing
    light in ingoing Kerr-Schild time).  Neither turns; the periapsis is the inner end."""
    T = lambda r: r + 4.0 * math.log(r - 2.0)  # noqa: E731
    for cap in (1.01, 1.005):
        v = look_view(0.0, 30.0, math.pi / 2, cap=cap)
        rep = report(v, [-1.0, 0.0, 0.0])
        assert rep["cls"] == O.CAPTURED
        assert rep["delay"] == pytest.approx(T(30.0) - T(2.0 * cap), rel=1e-9)
        assert rep["turns"] == pytest.approx(0.0, abs=1e-14)
        assert rep["r_min"] == pytest.approx(2.0 * cap, rel=1e-14)
        assert rep["b"] == 0.0 and rep["q"] == pytest.approx(0.0, abs=1e-20)
    for esc in (250.0, 1000.0):
        rep = report(look_view(0.0, 30.0, math.pi / 2, esc=esc), [1.0, 0.0, 0.0])
        assert rep["cls"] == O.ESCAPED
        assert rep["delay"] == pytest.approx(esc - 30.0, rel=1e-10)
        assert rep["r_min"] == 30.0 and abs(rep["turns"]) < 1e-14


@pytest.mark.parametrize("b", [5.5, 6.0, 8.0, 12.0, 20.0])
@pytest.mark.parametrize("incl", [math.pi / 2, 1.0])
def test_schwarzschild_turns_periapsis_and_delay(b, incl):
    """a = 0 rays of impact parameter b from r = 30 through their periapsis out to r_esc =
    250: turns, r_min and delay against the exact integrals.  At incl != pi/2 the same
    ray in a tilted plane (the look turned about the radial axis): the same answers."""
    r, esc = 30.0, 250.0
    lk = look_for(b, r)
    if incl != math.pi / 2:  # turn the look about e_r: the ray's plane tilts, nothing else changes
        c, s = math.cos(0.7), math.sin(0.7)
        lk = [lk[0], -s * lk[2], c * lk[2]]
    rep = report(look_view(0.0, r, incl, esc=esc), lk)
    assert rep["cls"] == O.ESCAPED
    ut = periapsis_u(b, r)
    assert rep["r_min"] == pytest.approx(1.0 / ut, rel=1e-12)
    sweep = through_periapsis(b, ut, r, lambda u: 1.0) + through_periapsis(b, ut, esc, lambda u: 1.0)
    assert rep["turns"] == pytest.approx(sweep / (2.0 * math.pi), abs=1e-10)
    f = lambda u: 1.0 / (b * u * u * (1.0 - 2.0 * u))  # noqa: E731
    t_bl = through_periapsis(b, ut, r, f) + through_periapsis(b, ut, esc, f)
    assert rep["delay"] == pytest.approx(t_bl + 2.0 * math.log((r - 2.0) / (esc - 2.0)), rel=1e-9)
    # looking toward +e_phi the light arrives moving toward -phi: L = -b sin(theta) times
    # the look's share along e_phi, and Q = b^2 - L^2 at a = 0
    bz = -b * math.sin(incl) * lk[2] / math.hypot(lk[1], lk[2])
    assert rep["b"] == pytest.approx(bz, rel=1e-12)
    assert rep["q"] == pytest.approx(b * b - bz * bz, rel=1e-9, abs=1e-12)


def test_turns_count_whole_orbits():
    """Near the critical impact parameter the sweep grows like -ln(b - b_c) / (2 pi)
    (the photon-sphere winding, Lyapunov exponent 1 at a = 0): each factor e^(2 pi) closer
    adds one turn."""
    bc, r = 3.0 * math.sqrt(3.0), 1000.0
    t = [report(look_view(0.0, r, math.pi / 2, esc=250.0), look_for(bc * (1.0 + eps), r))["turns"]
         for eps in (1e-4, 1e-4 * math.exp(-2.0 * math.pi))]
    assert t[1] - t[0] == pytest.approx(1.0, abs=2e-3)


def test_reference_traceray_claims():
    """What the reference's testRayTrace asserts of its RayReport (test_kerr.cpp:496-543),
    on the oracle: the grazing b_c (1 + 1e-6) ray circles more than 1.5 turns with
    r_min within 0.01 of 3, the 2 b_c ray turns less than 0.7."""
    bc, r = 3.0 * math.sqrt(3.0), 1000.0
    g1 = report(look_view(0.0, r, math.pi / 2), look_for(bc * (1.0 + 1e-6), r))
    g2 = report(look_view(0.0, r, math.pi / 2), look_for(2.0 * bc, r))
    assert g1["cls"] == O.ESCAPED and g1["turns"] > 1.5 and abs(g1["r_min"] - 3.0) < 0.01
    assert g2["cls"] == O.ESCAPED and g2["turns"] < 0.7


def test_reports_match_the_reference_at_its_accuracy():
    """The reference's own RayReport for 46 rays (its testRayTrace rays and seeded looks at
    a = 0, 0.5, 0.9, 0.998, with and without a disk), the capture radius set to its r+ + 0.01.
    The reference stops one RK4 step past each event and sums chords of its steps, so:
    the same end; turns within 1.5e-3 (3e-3 for disk hits, whose last chord runs past the
    disk); r_min within 1e-4 relative of the reference's (a step end, so above the true
    periapsis by up to its step), except where the reference's last step ran past the disk
    or below the capture radius; sky delays within 5e-4 relative once the reference's
    overshoot past r_esc is taken off, disk delays within 1e-4; b and q exact.  Delays at
    capture are not compared: the reference's step below r+ + 0.01 lands where dt/dr
    diverges as 1/(r - r+)."""
    n = 0
    for case, ref in zip(G["cases"], G["report"]):
        a, r, th, lx, ly, lz, disk, rin, rout, esc = case
        rep = report(look_view(a, r, th, int(disk), rin, rout, esc), [lx, ly, lz])
        end = int(ref[0])
        assert rep["cls"] == {0: O.ESCAPED, 1: O.CAPTURED, 2: O.DISK}[end], case
        assert rep["b"] == pytest.approx(ref[7], rel=1e-12, abs=1e-13)
        assert rep["q"] == pytest.approx(ref[8], rel=1e-11, abs=1e-12)
        assert abs(rep["turns"] - ref[3]) < (3e-3 if end == 2 else 1.5e-3), case
        r_esc = max(esc, 1.5 * r)
        if end == 0:
            assert rep["r_min"] == pytest.approx(ref[4], rel=1e-4)
            assert rep["r_min"] <= ref[4] + 1e-12
            assert rep["delay"] == pytest.approx(ref[5] - (ref[1] - r_esc), rel=5e-4)
        elif end == 2:
            # the reference's least r includes its step past the disk, so it may lie below the
            # disk radius; ours stops at the crossing
            assert ref[4] * (1.0 - 1e-4) <= rep["r_min"] <= rep["r_hit"] + 1e-12
            if rep["r_min"] < rep["r_hit"] - 1e-6:  # a periapsis before the disk
                assert rep["r_min"] == pytest.approx(ref[4], rel=1e-4)
            assert rep["delay"] == pytest.approx(ref[5], rel=1e-4)
        else:
            assert rep["r_min"] == pytest.approx((1.0 + 0.01 / L.ko_horizon_radius(a)) * L.ko_horizon_radius(a),
                                                 rel=1e-14)
            assert ref[4] <= rep["r_min"]
        n += 1
    assert n == 46
