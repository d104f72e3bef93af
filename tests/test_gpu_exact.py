"""The GPU integrator against exact answers (tests/quadratures.py), not only against the
oracle: sky angles of equatorial rays (a = 0 and a = 0.9, prograde and retrograde) and
disk-hit radii with their g (a = 0 over the pole, a = 0.9 from i = 60 deg).  The FP64 path
holds 1e-9 rad / 1e-9 relative (measured 1.4e-10 and 3e-11 at worst); the FP32/FP64 path
1e-5 (measured 1.0e-6 and 4.4e-7: its bulk runs in FP32)."""
import math

import numpy as np
import pytest
import quadratures as Qd

import black_simulation_amd as bsa

pytestmark = pytest.mark.gpu

BOUND = {bsa.PREC_FP64: 1e-9, bsa.PREC_MIXED: 1e-5}
G_BOUND = {bsa.PREC_FP64: 1e-12, bsa.PREC_MIXED: 1e-6}
PRECISIONS = [bsa.PREC_FP64, bsa.PREC_MIXED]


def wrapped(d):
    return (d + math.pi) % (2 * math.pi) - math.pi


@pytest.mark.parametrize("prec", PRECISIONS, ids=["fp64", "mixed"])
def test_equatorial_sky_angles_exact(tracer, prec):
    """edge-on cameras at r = 50 M, beta = 0: phi at infinity against the orbit quadrature"""
    cases = [(0.0, b, Qd.schwarzschild_sweep(b, 50.0)) for b in (6.0, 7.5, 10.0, 20.0)]
    cases = [(a, al, (i + o) % (2 * math.pi)) for a, al, (i, o, _) in cases]
    cases += [(0.9, al, Qd.kerr_equatorial_phi(0.9, al, 50.0)) for al in (-4.0, -8.0, 8.0, 12.0)]
    for a, al, want in cases:
        v = bsa.View(spin=a, r_cam=50.0, theta_cam=math.pi / 2, disk_enabled=0)
        print(f"equatorial a={a} alpha={al} prec={prec}", flush=True)
        got = bsa.to_numpy(tracer.trace_params(v, [[al, 0.0]], bsa.TraceOpts(precision=prec)))
        assert got["cls"][0] & 7 == bsa.ESCAPED
        d = abs(wrapped(got["phi_inf"][0] - want))
        print(f"  phi error {d:.2e} rad", flush=True)
        assert d < BOUND[prec], (a, al, got["phi_inf"][0], want)
        assert abs(got["theta_inf"][0] - math.pi / 2) < BOUND[prec]


@pytest.mark.parametrize("prec", PRECISIONS, ids=["fp64", "mixed"])
def test_disk_hits_exact(tracer, prec):
    """r_hit against the quadratures; g against the Keplerian closed form at the traced r"""
    rows = []
    for b in (5.4, 5.7, 6.2):
        v = bsa.View(spin=0.0, r_cam=30.0, theta_cam=math.pi / 2, disk_inner=3.05, disk_outer=40.0)
        rows.append((v, 0.0, b, Qd.schwarzschild_pole_disk_hit(b, 30.0), 0.0))
    inc = math.radians(60.0)
    for al, be in ((0.0, -3.0), (3.0, -4.0), (-5.0, -2.0), (2.0, -7.0), (6.0, -1.5), (-8.0, -3.0)):
        v = bsa.View(spin=0.9, r_cam=50.0, theta_cam=inc, disk_inner=1.5, disk_outer=100.0)
        rows.append((v, al, be, Qd.kerr_disk_hit(0.9, 50.0, inc, al, be), -al * math.sin(inc)))
    for v, al, be, r_want, Lz in rows:
        print(f"disk a={v.spin} alpha={al} beta={be} prec={prec}", flush=True)
        got = bsa.to_numpy(tracer.trace_params(v, [[al, be]], bsa.TraceOpts(precision=prec)))
        assert got["cls"][0] & 7 == bsa.DISK
        r = got["r_hit"][0]
        print(f"  r_hit rel error {abs(r / r_want - 1):.2e}", flush=True)
        assert r == pytest.approx(r_want, rel=BOUND[prec])
        assert got["g"][0] == pytest.approx(Qd.keplerian_g(v.spin, r, Lz), rel=G_BOUND[prec])
        assert np.isnan(got["theta_inf"][0])
