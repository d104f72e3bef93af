"""Batched uneven views in one launch, the hipGraph plan and the autotuner."""
import numpy as np
import pytest
import torch

import black_simulation_amd as bsa
from black_simulation_amd.configs import c1_view, c2_view, c3_views
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_batch():
    views = c3_views(n=6, sizes=(16, 24, 32))
    views.append(bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=20, height=12, spin=0.7, r_cam=25.0, theta_cam=1.3))
    return views


def test_seeded_batch_matches_oracle_per_view(tracer):
    """Views of different size, spin, inclination and camera radius share one work queue."""
    views = small_batch()
    got = bsa.to_numpy(tracer.trace_views(views))
    off = 0
    for i, v in enumerate(views):
        n = v.rays
        ov = O.make_view(**{k: getattr(v, k) for k in (
            "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width",
            "half_height", "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde",
            "disk_enabled", "escape_radius", "capture_factor")})
        want = O.render(ov, O.make_opts())
        assert np.array_equal(got["cls"][off:off + n], want["cls"]), f"view {i}"
        assert np.array_equal(got["steps"][off:off + n], want["steps"]), f"view {i}"
        off += n


def test_plan_replays_equal_eager(tracer):
    """The hipGraph plan, replayed several times with eager work and synchronisation in
    between, reproduces the eager launch every time."""
    views = small_batch()
    eager = bsa.to_numpy(tracer.trace_views(views))
    plan = tracer.plan(views)
    for rep in range(3):
        out = plan.launch()
        torch.cuda.synchronize()
        got = bsa.to_numpy(out)
        for k in ("cls", "steps", "r_hit", "g", "theta_inf"):
            assert np.array_equal(got[k], eager[k], equal_nan=True), (rep, k)
        tracer.trace_views([c1_view()])  # unrelated eager work between replays
        torch.cuda.synchronize()
        for k in out:
            out[k].zero_()
    plan.close()


def test_autotune_returns_a_valid_shape_and_results_do_not_depend_on_it(tracer):
    v = c2_view(192, 108)
    tuned = tracer.autotune([v])
    assert tuned.block in (64, 128, 256) and 1 <= tuned.waves_per_simd <= 4
    assert tuned.waves_per_cu > 0 and tuned.steps_per_dequeue > 0
    print("autotuned", tuned, flush=True)
    a = bsa.to_numpy(tracer.trace_views([v], tuned))
    b = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(mode=bsa.MODE_ONESHOT)))
    for k in ("cls", "steps", "r_hit", "g"):
        assert np.array_equal(a[k], b[k], equal_nan=True)
