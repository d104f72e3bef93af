"""Edge cases of the FP64 path against the oracle, bit for bit: the extremal and a
retrograde spin, a camera on the spin axis (theta clamped to 1e-4, as eht.cpp:266-267),
a step budget that runs out (class unfinished, as trace.frag:1796-1800), one-pixel and
one-row views; and the boundary's handling of empty calls."""
import ctypes as C
import math

import numpy as np
import pytest

import black_simulation_amd as bsa
from black_simulation_amd import _lib as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_view(v: bsa.View):
    return O.make_view(**{k: getattr(v, k) for k in (
        "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
        "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
        "escape_radius", "capture_factor")})


def same(tracer, v, opts=None, oopts=None, what=""):
    got = bsa.to_numpy(tracer.trace_views([v], opts or bsa.TraceOpts()))
    want = O.render(oracle_view(v), oopts or O.make_opts())
    print(f"{what}: classes {np.bincount(want['cls'], minlength=4)}", flush=True)
    assert np.array_equal(got["cls"] & 7, want["cls"]), what
    assert np.array_equal(got["steps"], want["steps"]), what
    esc = want["cls"] == O.ESCAPED
    if esc.any():
        assert np.max(np.abs(got["theta_inf"][esc] - want["theta"][esc])) <= 1e-13, what
    disk = want["cls"] == O.DISK
    assert np.array_equal(got["r_hit"][disk], want["r_hit"][disk]), what
    assert np.array_equal(got["g"][disk], want["g"][disk]), what
    return want


@pytest.mark.parametrize("spin", [1.0, -0.9, 0.0])
def test_extremal_retrograde_and_schwarzschild_spins(tracer, spin):
    v = bsa.View(width=32, height=32, spin=spin, r_cam=30.0, theta_cam=math.radians(70.0), half_width=10.0,
                 half_height=10.0, disk_outer=20.0)
    want = same(tracer, v, what=f"a={spin}")
    assert (want["cls"] == O.CAPTURED).any() and (want["cls"] == O.DISK).any()


def test_camera_on_the_spin_axis(tracer):
    v = bsa.View(width=24, height=24, spin=0.9, r_cam=30.0, theta_cam=0.0, half_width=10.0, half_height=10.0)
    same(tracer, v, what="theta_cam = 0")


def test_step_budget_runs_out(tracer):
    v = bsa.View(width=24, height=24, spin=0.9, r_cam=30.0, theta_cam=math.radians(80.0), half_width=10.0,
                 half_height=10.0)
    want = same(tracer, v, bsa.TraceOpts(max_steps=40), O.make_opts(max_steps=40), what="max_steps 40")
    assert (want["cls"] == 3).any() and (want["steps"] <= 40).all()


def test_one_pixel_and_one_row_views(tracer):
    same(tracer, bsa.View(width=1, height=1, spin=0.9, r_cam=30.0, half_width=0.5, half_height=0.5), what="1x1")
    same(tracer, bsa.View(width=200, height=1, spin=0.9, r_cam=30.0, half_width=12.0, half_height=0.06), what="200x1")


def test_non_finite_view_fields_are_refused(tracer):
    """NaN or infinity in any real field of kt_view is KT_EINVAL with a message, also in the
    fields no range check covers (phi, yaw, the observer's frame)"""
    lib = K.lib()
    o = bsa.TraceOpts().to_c()
    out = K.KtRaysOut()
    for field, value in (("phi_cam", math.nan), ("yaw", math.inf), ("spin", math.nan), ("r_cam", math.inf),
                         ("escape_radius", math.nan), ("disk_inner", -math.inf)):
        v = bsa.View(width=4, height=4)
        setattr(v, field, value)
        assert lib.kt_trace_views(tracer._ctx, C.byref(v.to_c()), 1, C.byref(o), C.byref(out), None) == K.KT_EINVAL
        assert b"finite" in lib.kt_last_error_string(tracer._ctx), field
    fr = [float(i % 5 == 0) for i in range(16)]
    fr[5] = math.nan
    v = bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=4, height=4, spin=0.9, r_cam=10.0, theta_cam=1.2,
                 frame=tuple(fr))
    assert lib.kt_trace_views(tracer._ctx, C.byref(v.to_c()), 1, C.byref(o), C.byref(out), None) == K.KT_EINVAL
    assert b"frame must be finite" in lib.kt_last_error_string(tracer._ctx)


def test_empty_calls_are_refused_or_do_nothing(tracer):
    lib = K.lib()
    out = K.KtRaysOut()
    v = bsa.View().to_c()
    o = bsa.TraceOpts().to_c()
    assert lib.kt_trace_views(tracer._ctx, C.byref(v), 0, C.byref(o), C.byref(out), None) == K.KT_EINVAL
    assert lib.kt_trace_rays(tracer._ctx, C.byref(v), None, 0, C.byref(o), C.byref(out), None) == K.KT_EINVAL
    bad = bsa.View(width=0).to_c()
    assert lib.kt_trace_views(tracer._ctx, C.byref(bad), 1, C.byref(o), C.byref(out), None) == K.KT_EINVAL
    cols = [np.zeros(1) for _ in range(8)] + [np.zeros(1, np.int8)]
    s = K.KtParticles(*[c.ctypes.data for c in cols], None, None, 0)
    assert lib.kt_particles_advance(tracer._ctx, 0.9, C.byref(s), 0, 10.0, 0.0, 1e4, C.byref(o), None) == K.KT_OK
    assert lib.kt_particles_advance(tracer._ctx, 0.9, C.byref(K.KtParticles()), 1, 10.0, 0.0, 1e4, C.byref(o),
                                    None) == K.KT_EINVAL
