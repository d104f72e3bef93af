"""FP64 GPU path against the FP64 CPU oracle on identical rays.

The step loop performs only correctly rounded IEEE operations in the same order on both
sides, so classes, step counts, disk radii and g must be identical bit for bit; the sky
angles pass through acos (device libm vs host libm) once at the end, so they are held
to 1e-13 rad.  Bar from the task: class >= 99.9 %, deflection <= 1e-9 rad, g <= 1e-10 rel.
"""
import math

import numpy as np
import pytest

import black_simulation_amd as bsa
from black_simulation_amd.configs import c1_view, c2_view
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ANGLE_TOL = 1e-13


def oracle_view(v: bsa.View):
    return O.make_view(**{k: getattr(v, k) for k in (
        "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
        "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
        "escape_radius", "capture_factor")})


def sky_angle(t1, p1, t2, p2):
    """angle between two sky directions (theta, phi)"""
    c = np.sin(t1) * np.sin(t2) * np.cos(p1 - p2) + np.cos(t1) * np.cos(t2)
    s = np.sqrt(np.maximum(0.0, (np.sin(t1) * np.cos(p1) - np.sin(t2) * np.cos(p2)) ** 2 +
                           (np.sin(t1) * np.sin(p1) - np.sin(t2) * np.sin(p2)) ** 2 + (np.cos(t1) - np.cos(t2)) ** 2))
    return 2.0 * np.arcsin(np.minimum(1.0, s / 2.0))


def compare(got, want, tag):
    print(f"[{tag}] classes gpu={np.bincount(got['cls'] & 7, minlength=4)} "
          f"oracle={np.bincount(want['cls'], minlength=4)}", flush=True)
    assert np.array_equal(got["cls"] & 7, want["cls"]), f"{tag}: class mismatch"
    assert np.array_equal(got["steps"], want["steps"]), f"{tag}: step counts differ"
    esc = want["cls"] == O.ESCAPED
    if esc.any():
        d = sky_angle(got["theta_inf"][esc], got["phi_inf"][esc], want["theta"][esc], want["phi"][esc])
        print(f"[{tag}] max sky-direction difference {d.max():.3e} rad over {esc.sum()} escaped rays", flush=True)
        assert d.max() <= ANGLE_TOL
    disk = want["cls"] == O.DISK
    if disk.any():
        assert np.array_equal(got["r_hit"][disk], want["r_hit"][disk]), f"{tag}: r_hit differs"
        assert np.array_equal(got["g"][disk], want["g"][disk]), f"{tag}: g differs"
    assert np.all(np.isnan(got["theta_inf"][~esc]))
    assert np.all(np.isnan(got["g"][~disk]))


@pytest.mark.parametrize("mode", [bsa.MODE_PERSISTENT, bsa.MODE_ONESHOT])
def test_c1_small_bit_exact(tracer, mode):
    v = c1_view()
    v.width = v.height = 64
    print(f"c1 64x64 mode={mode}", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(mode=mode)))
    want = O.render(oracle_view(v), O.make_opts())
    compare(got, want, f"c1-64 mode{mode}")


def test_c1_full_bit_exact(tracer):
    """The CPU-runnable configuration in full: 256x256, a=0.9, r=30, i=80 deg."""
    v = c1_view()
    print("c1 256x256", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v]))
    want = O.render(oracle_view(v), O.make_opts())
    compare(got, want, "c1-full")


def test_c2_sample_bit_exact(tracer):
    """a=0.998, i=85 deg, 1080p plane: a 96x54 subsampling of the same 30 M-wide plane."""
    v = c2_view(96, 54)
    print("c2 96x54", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v]))
    want = O.render(oracle_view(v), O.make_opts())
    compare(got, want, "c2-96x54")


def test_pinhole_bit_exact(tracer):
    """The simulator's own camera (ZAMO frame, yaw/pitch/roll, fov) at 48x32."""
    v = bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=48, height=32, spin=0.94, r_cam=32.0, theta_cam=1.48,
                 yaw=0.05, pitch=-0.02, roll=0.1, fov_deg=60.0, disk_outer=20.0)
    print("pinhole 48x32", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v]))
    want = O.render(oracle_view(v), O.make_opts())
    compare(got, want, "pinhole")


def test_batched_views_bit_exact(tracer):
    """Several views of different size, spin and camera kind in one launch."""
    views = [c1_view(), c2_view(64, 36),
             bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=24, height=16, spin=0.5, r_cam=20.0, theta_cam=1.2)]
    views[0].width = views[0].height = 40
    print("batch of 3 views", flush=True)
    got = bsa.to_numpy(tracer.trace_views(views))
    off = 0
    for i, v in enumerate(views):
        n = v.rays
        want = O.render(oracle_view(v), O.make_opts())
        compare({k: a[off:off + n] for k, a in got.items()}, want, f"batch-view{i}")
        off += n


def test_explicit_rays(tracer):
    """kt_trace_rays with explicit (alpha, beta), as eht::traceRay."""
    v = c1_view()
    rng = np.random.default_rng(5)
    ab = rng.uniform(-12, 12, size=(200, 2))
    got = bsa.to_numpy(tracer.trace_params(v, ab))
    ov = oracle_view(v)
    opts = O.make_opts()
    for i, (al, be) in enumerate(ab):
        p = O.Photon()
        y0 = (O.C.c_double * 5)()
        O.lib().ko_launch_bardeen(O.C.byref(ov), al, be, O.C.byref(p), y0)
        res = O.Result()
        O.lib().ko_trace_photon(O.C.byref(ov), O.C.byref(opts), O.C.byref(p), y0, O.C.byref(res))
        assert got["cls"][i] == res.cls and got["steps"][i] == res.steps
        if res.cls == O.ESCAPED:
            assert abs(got["theta_inf"][i] - res.theta_inf) <= ANGLE_TOL


def test_host_pointer_outputs():
    """on_device = 0: the library stages and copies back itself."""
    import ctypes as C

    from black_simulation_amd import _lib as K
    L = K.lib()
    ctx = C.c_void_p()
    assert L.kt_create(C.byref(K.KtDeviceOpts(0)), C.byref(ctx)) == 0
    v = c1_view()
    v.width = v.height = 16
    n = v.rays
    cls = np.zeros(n, np.int8)
    th = np.zeros(n)
    steps = np.zeros(n, np.int32)
    out = K.KtRaysOut(cls.ctypes.data, th.ctypes.data, None, None, None, steps.ctypes.data, 0)
    opts = K.KtTraceOpts()
    L.kt_trace_opts_default(C.byref(opts))
    rc = L.kt_trace_views(ctx, C.byref(v.to_c()), 1, C.byref(opts), C.byref(out), None)
    assert rc == 0, L.kt_last_error_string(ctx)
    want = O.render(oracle_view(v), O.make_opts())
    assert np.array_equal(cls, want["cls"]) and np.array_equal(steps, want["steps"])
    L.kt_destroy(ctx)


def test_pinned_host_buffers_reused_across_calls():
    """trace_views_host into page-locked buffers allocated once (host_outputs) gives, call
    after call, what device outputs give; fresh numpy buffers give the same."""
    tr = bsa.Tracer(0)
    v = c1_view()
    v.width = v.height = 48
    hb = tr.host_outputs(v.rays)
    assert hb["theta_inf"].is_pinned()
    dev = bsa.to_numpy(tr.trace_views([v]))
    for _ in range(2):
        got = tr.trace_views_host([v], out=hb)
        for k, x in got.items():
            assert np.array_equal(x.numpy(), dev[k], equal_nan=True), k
    fresh = tr.trace_views_host([v])
    for k, x in fresh.items():
        assert np.array_equal(x, dev[k], equal_nan=True), k
    tr.close()


def test_float_outputs_are_the_doubles_rounded_once():
    """kt_rays_out's FP32 fields: on the device, through host buffers and from a plan, each
    is the FP64 output rounded to float (NaN where the output does not apply); the surrogate
    path's answered rays too."""
    import torch

    tr = bsa.Tracer(0)
    v = c2_view(96, 54)
    want = bsa.to_numpy(tr.trace_views([v]))
    f = bsa.to_numpy(tr.trace_views([v], out=tr.alloc_outputs(v.rays, f32=True)))
    h = tr.trace_views_host([v], out=tr.host_outputs(v.rays, f32=True))
    plan = tr.plan([v], out=tr.alloc_outputs(v.rays, f32=True))
    p = plan.launch()
    torch.cuda.synchronize()
    p = bsa.to_numpy(p)
    plan.close()
    for got in (f, {k: x.numpy() for k, x in h.items()}, p):
        assert np.array_equal(got["cls"], want["cls"]) and np.array_equal(got["steps"], want["steps"])
        for k in ("theta_inf", "phi_inf", "r_hit", "g"):
            assert got[k].dtype == np.float32
            assert np.array_equal(got[k], want[k].astype(np.float32), equal_nan=True), k
    from black_simulation_amd.train_surrogate import load_params

    tr.surrogate_load(load_params())
    opts = bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED)
    from black_simulation_amd.configs import c5_view

    sv = c5_view(30, 96, 54)
    d = bsa.to_numpy(tr.trace_views([sv], opts))
    s = bsa.to_numpy(tr.trace_views([sv], opts, out=tr.alloc_outputs(sv.rays, f32=True)))
    assert ((d["cls"] & bsa.FLAG_SURROGATE) != 0).any()
    assert np.array_equal(s["cls"], d["cls"])
    for k in ("theta_inf", "phi_inf", "r_hit", "g"):
        assert np.array_equal(s[k], d[k].astype(np.float32), equal_nan=True), k
    tr.close()
