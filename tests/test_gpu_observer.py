"""Observer cameras on the GPU (KT_CAMERA_OBSERVER: boosted ZAMOs, falling cameras inside
and outside the horizon) against the oracle on the same rays.  The FP64 integrator with the
Kerr-Schild phi performs the oracle's operations in the oracle's order, so classes and step
counts are equal and the outputs agree to the last bits (1e-12 relative); the far-side
rule gives the same dark rays.  The existing camera kinds are untouched by the observer
kernel (their parity tests still hold bit for bit)."""
import math

import numpy as np
import pytest
import torch

import black_simulation_amd as bsa
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OL = O.lib()


def both(v: bsa.View, opts=None):
    tr = bsa.Tracer(0)
    out = tr.trace_views([v], opts or bsa.TraceOpts(precision=bsa.PREC_FP64))
    torch.cuda.synchronize()
    got = bsa.to_numpy(out)
    tr.close()
    ov = O.make_view(**{k: getattr(v, k) for k in (
        "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
        "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
        "escape_radius", "capture_factor", "u_obs", "drift", "frame")})
    want = O.render(ov, O.make_opts(), threads=16)
    return got, want


def sky_angle(t1, p1, t2, p2):
    s = np.sqrt((np.sin(t1) * np.cos(p1) - np.sin(t2) * np.cos(p2)) ** 2 +
                (np.sin(t1) * np.sin(p1) - np.sin(t2) * np.sin(p2)) ** 2 + (np.cos(t1) - np.cos(t2)) ** 2)
    return 2.0 * np.arcsin(np.minimum(1.0, s / 2.0))


def assert_parity(got, want, what):
    cls = got["cls"] & 7
    print(f"{what}: {len(cls)} rays, classes gpu {np.bincount(cls, minlength=4)} "
          f"oracle {np.bincount(want['cls'], minlength=4)}", flush=True)
    assert np.array_equal(cls, want["cls"]), what
    assert np.array_equal(got["steps"], want["steps"]), what
    esc = want["cls"] == O.ESCAPED
    if esc.any():
        # acos at the end and the phi_KS -> phi_BL log1p run on device and host libm
        d = sky_angle(got["theta_inf"][esc], got["phi_inf"][esc], want["theta"][esc], want["phi"][esc])
        assert d.max() <= 1e-12, (what, d.max())
    disk = want["cls"] == O.DISK
    assert np.array_equal(got["r_hit"][disk], want["r_hit"][disk]), what
    assert np.array_equal(got["g"][disk], want["g"][disk]), what


def falling_u(a, r, th, E=1.0, L=0.0, Q=0.0):
    return bsa.geodesic_four_velocity(a, E, L, Q, r, th, -1, 1)


@pytest.mark.parametrize("case", ["boost_phi", "boost_mixed", "falling_inside", "falling_outside", "schwarzschild_inside"])
def test_observer_views_match_the_oracle(case):
    a, r, th, u, drift = 0.9, 10.0, 1.2, (0, 0, 0, 0), (0, 0, 0)
    if case == "boost_phi":
        drift = (0.0, 0.0, 0.6)
    elif case == "boost_mixed":
        drift = (0.3, -0.2, 0.4)
    elif case == "falling_inside":
        r, th = 1.2, 1.0
        u = falling_u(a, r, th, 1.0, 2.0, 3.0)
    elif case == "falling_outside":
        r, th = 1.6, 1.0
        u = falling_u(a, r, th)
    else:
        a, r, th = 0.0, 1.5, 1.4
        u = falling_u(a, r, th)
    v = bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=48, height=32, spin=a, r_cam=r, theta_cam=th, phi_cam=0.4,
                 yaw=0.3, pitch=-0.2, fov_deg=100.0, disk_inner=0.0, disk_outer=20.0, disk_enabled=1,
                 escape_radius=1000.0, capture_factor=1.01, u_obs=u, drift=drift)
    got, want = both(v)
    assert_parity(got, want, case)
    if case == "falling_inside":
        assert (want["cls"] == O.CAPTURED).any() and (want["steps"][want["cls"] == O.CAPTURED] == 0).any()


def test_observer_explicit_looks_and_plans():
    """kt_trace_rays with look vectors and a replayed plan give the same rays as the view."""
    a, r, th = 0.9, 1.3, 1.1
    u = falling_u(a, r, th)
    v = bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=16, height=16, spin=a, r_cam=r, theta_cam=th,
                 fov_deg=90.0, disk_enabled=0, capture_factor=1.01, u_obs=u)
    tr = bsa.Tracer(0)
    opts = bsa.TraceOpts(precision=bsa.PREC_FP64)
    ref = bsa.to_numpy(tr.trace_views([v], opts))
    plan = tr.plan([v], opts)
    for _ in range(3):
        out = plan.launch()
        torch.cuda.synchronize()
        got = bsa.to_numpy(out)
        assert np.array_equal(got["cls"], ref["cls"]) and np.array_equal(got["steps"], ref["steps"])
        m = np.isfinite(ref["theta_inf"])
        assert np.array_equal(got["theta_inf"][m], ref["theta_inf"][m])
    plan.close()
    looks = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [0.3, 0.9, 0.3]])
    looks /= np.linalg.norm(looks, axis=1, keepdims=True)
    out = bsa.to_numpy(tr.trace_params(v, looks, opts))
    want = O.trace_params(O.make_view(camera_kind=O.OBSERVER, width=1, height=1, spin=a, r_cam=r, theta_cam=th,
                                      disk_enabled=0, escape_radius=1000.0, capture_factor=1.01, u_obs=u),
                          O.make_opts(), looks)
    assert np.array_equal(out["cls"] & 7, want["cls"]) and np.array_equal(out["steps"], want["steps"])
    tr.close()


def test_observer_batch_rules():
    tr = bsa.Tracer(0)
    obs = bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=4, height=4, r_cam=10.0, theta_cam=1.2, drift=(0, 0, 0.5))
    zamo = bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=4, height=4, r_cam=10.0, theta_cam=1.2)
    with pytest.raises(RuntimeError, match="share a call"):
        tr.trace_views([obs, zamo], bsa.TraceOpts(precision=bsa.PREC_FP64))
    with pytest.raises(RuntimeError, match="FP64"):
        tr.trace_views([obs], bsa.TraceOpts(precision=bsa.PREC_MIXED))
    inside = bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=4, height=4, spin=0.9, r_cam=1.2, theta_cam=1.0)
    with pytest.raises(RuntimeError, match="hovers"):
        tr.trace_views([inside], bsa.TraceOpts(precision=bsa.PREC_FP64))
    tr.close()


def test_explicit_observer_frame():
    """A view given its frame outright (kt_view.frame, as uCamPhoton carries Camera::frame):
    the frame kt_observer_frame builds from u_obs traces exactly as u_obs does, a frame with
    its axes turned matches the oracle given the same frame, and a frame that is not
    orthonormal is refused."""
    a, r, th = 0.9, 1.2, 1.0
    u = falling_u(a, r, th, 1.0, 2.0, 3.0)
    base = dict(camera_kind=bsa.CAMERA_OBSERVER, width=48, height=32, spin=a, r_cam=r, theta_cam=th, phi_cam=0.4,
                yaw=0.3, pitch=-0.2, fov_deg=100.0, disk_inner=0.0, disk_outer=20.0, disk_enabled=1,
                escape_radius=1000.0, capture_factor=1.01)
    e = bsa.observer_frame(a, r, th, u)
    by_u, _ = both(bsa.View(**base, u_obs=u))
    got, want = both(bsa.View(**base, frame=tuple(e.ravel())))
    for k in ("cls", "steps", "theta_inf", "phi_inf", "r_hit", "g"):
        assert np.array_equal(got[k], by_u[k], equal_nan=True), k
    assert_parity(got, want, "frame from u_obs")
    c, s = math.cos(0.5), math.sin(0.5)
    e2 = e.copy()
    e2[1:] = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]]) @ e[1:]
    got, want = both(bsa.View(**base, frame=tuple(e2.ravel())))
    assert_parity(got, want, "turned frame")
    e3 = e.copy()
    e3[2] *= 1.001
    tr = bsa.Tracer(0)
    with pytest.raises(RuntimeError, match="orthonormal"):
        tr.trace_views([bsa.View(**base, frame=tuple(e3.ravel()))], bsa.TraceOpts(precision=bsa.PREC_FP64))
    tr.close()
