"""The single-ray report on the GPU (kt_rays_out turns, r_min, delay, b, q; the RayReport of
trajectory::traceRay, src/trajectory.h:97-113) against the oracle's ko_trace_photon_report.

The report kernel integrates as the FP64 kernels do, operation for operation, so the
classes, step counts, disk radii and g are equal to the oracle's and to a trace without the
report; the report sums sines, cosines, logarithms and arctangents from the device's and
the host's maths libraries, so turns and delay agree to 1e-12 relative, r_min to 1e-13."""
import math
import os

import numpy as np
import pytest
import torch

import black_simulation_amd as bsa
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OL = O.lib()
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "report_vectors.npz"))
FIELDS = ("camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
          "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
          "escape_radius", "capture_factor", "u_obs", "drift", "frame")


def oracle_view(v):
    return O.make_view(**{k: getattr(v, k) for k in FIELDS})


def assert_report(got, want, what):
    cls = got["cls"] & 7
    assert np.array_equal(cls, want["cls"]), what
    assert np.array_equal(got["steps"], want["steps"]), what
    disk = want["cls"] == O.DISK
    assert np.array_equal(got["r_hit"][disk], want["r_hit"][disk]), what
    assert np.array_equal(got["g"][disk], want["g"][disk]), what
    np.testing.assert_allclose(got["turns"], want["turns"], rtol=1e-12, atol=1e-13, err_msg=what)
    np.testing.assert_allclose(got["delay"], want["delay"], rtol=1e-12, atol=1e-12, err_msg=what)
    np.testing.assert_allclose(got["r_min"], want["r_min"], rtol=1e-13, err_msg=what)
    assert np.array_equal(got["b"], want["b"]) and np.array_equal(got["q"], want["q"]), what


def test_report_of_zamo_looks_matches_oracle(tracer):
    """The 46 rays of the reference's report fixture (its testRayTrace rays and seeded looks
    at four spins, with and without a disk), each through kt_trace_rays with the report."""
    opts = bsa.TraceOpts(precision=bsa.PREC_FP64)
    for k, case in enumerate(G["cases"]):
        a, r, th, lx, ly, lz, disk, rin, rout, esc = case
        rh = OL.ko_horizon_radius(a)
        v = bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=1, height=1, spin=a, r_cam=r, theta_cam=th, phi_cam=0.0,
                     disk_enabled=int(disk), disk_inner=rin, disk_outer=rout, escape_radius=esc,
                     capture_factor=1.0 + 0.01 / rh)
        got = bsa.to_numpy(tracer.trace_params(v, [[lx, ly, lz]], opts, report=True))
        want = O.trace_params(oracle_view(v), O.make_opts(), [[lx, ly, lz]], report=True)
        assert_report(got, want, f"case {k}")


@pytest.mark.parametrize("a,incl", [(0.0, 80.0), (0.9, 80.0), (0.998, 17.0)])
def test_report_of_a_whole_view(tracer, a, incl):
    """A 40 x 40 Bardeen view with a disk in one launch with the report: every ray equals
    the oracle's; the report's classes, steps, sky angles and disk hits are those of the
    same view traced without it."""
    v = bsa.View(width=40, height=40, spin=a, r_cam=30.0, theta_cam=math.radians(incl), half_width=12.0,
                 half_height=12.0, disk_inner=0.0, disk_outer=20.0, escape_radius=1000.0, capture_factor=1.01)
    opts = bsa.TraceOpts(precision=bsa.PREC_FP64)
    got = bsa.to_numpy(tracer.trace_views([v], opts, report=True))
    plain = bsa.to_numpy(tracer.trace_views([v], opts))
    for k in ("cls", "steps", "theta_inf", "phi_inf", "r_hit", "g"):
        assert np.array_equal(got[k], plain[k], equal_nan=True), k
    idx = np.arange(v.rays)
    al = -v.half_width + (idx % v.width + 0.5) * (2 * v.half_width / v.width)
    be = v.half_height - (idx // v.width + 0.5) * (2 * v.half_height / v.height)
    want = O.trace_params(oracle_view(v), O.make_opts(), np.stack([al, be], 1), report=True)
    assert_report(got, want, f"a={a}")
    print(f"a={a} i={incl}: turns max {got['turns'].max():.3f}, r_min min {got['r_min'].min():.4f}, "
          f"delay range {got['delay'].min():.2f}..{got['delay'].max():.2f}", flush=True)


def test_report_of_observer_cameras(tracer):
    """Observer cameras (Kerr-Schild phi): a boosted ZAMO and a camera falling inside the
    horizon, whose far-side rays end at once with delay 0 and r_min = r_cam."""
    for a, r, th, u, drift in [(0.9, 10.0, 1.2, (0, 0, 0, 0), (0.3, -0.2, 0.4)),
                               (0.9, 1.2, 1.0, tuple(bsa.geodesic_four_velocity(0.9, 1.0, 2.0, 3.0, 1.2, 1.0, -1, 1)),
                                (0, 0, 0))]:
        v = bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=24, height=16, spin=a, r_cam=r, theta_cam=th,
                     phi_cam=0.4, yaw=0.3, pitch=-0.2, fov_deg=100.0, disk_inner=0.0, disk_outer=20.0,
                     disk_enabled=1, escape_radius=1000.0, capture_factor=1.01, u_obs=u, drift=drift)
        opts = bsa.TraceOpts(precision=bsa.PREC_FP64)
        got = bsa.to_numpy(tracer.trace_views([v], opts, report=True))
        ov = oracle_view(v)
        want = {k: [] for k in ("cls", "steps", "r_hit", "g", "turns", "r_min", "delay", "b", "q")}
        for row in range(v.height):
            for col in range(v.width):
                p = O.Photon()
                y0 = (O.C.c_double * 5)()
                OL.ko_launch_pixel(O.C.byref(ov), col, row, O.C.byref(p), y0)
                res, rep = O.Result(), O.Report()
                OL.ko_trace_photon_report(O.C.byref(ov), O.C.byref(O.make_opts()), O.C.byref(p), y0, O.C.byref(res),
                                          O.C.byref(rep))
                want["cls"].append(res.cls)
                want["steps"].append(res.steps)
                want["r_hit"].append(res.r_hit if res.cls == O.DISK else np.nan)
                want["g"].append(res.g if res.cls == O.DISK else np.nan)
                for k in ("turns", "r_min", "delay", "b", "q"):
                    want[k].append(getattr(rep, k))
        want = {k: np.array(x) for k, x in want.items()}
        assert_report(got, want, f"observer r={r}")
        dark = (want["steps"] == 0) & (want["cls"] == O.CAPTURED)
        if r < 2:
            assert dark.any() and np.all(got["delay"][dark] == 0.0) and np.all(got["r_min"][dark] == r)


def test_report_rules(tracer):
    """The report is an FP64 trace: asked with the mixed path, the surrogate, the
    allocation-free entry or a plan, the call is refused."""
    v = bsa.View(width=4, height=4)
    with pytest.raises(RuntimeError, match="FP64"):
        tracer.trace_views([v], bsa.TraceOpts(precision=bsa.PREC_MIXED), report=True)
    out = tracer.alloc_outputs(16, report=True)
    with pytest.raises(RuntimeError, match="report"):
        tracer.plan([v], bsa.TraceOpts(), out=out)
