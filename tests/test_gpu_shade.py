"""HDR colour on the GPU (kt_shade_views) against the oracle's ko_shade_ray on the same
traced rays: the kernel computes in FP32 and stores IEEE halves, so the bar is half
precision (relative 2e-3 plus 1e-4 absolute), with grid-line and quadrant borders (where
FP32 and FP64 floor() may land in different cells) allowed for a few rays."""
import numpy as np
import pytest
import torch

import black_simulation_amd as bsa
from black_simulation_amd.configs import c5_view
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def traced(tracer, v):
    out = tracer.trace_views([v], bsa.TraceOpts(precision=bsa.PREC_FP64))
    torch.cuda.synchronize()
    return out, bsa.to_numpy(out)


def compare(got, want, name):
    tol = 2e-3 * np.abs(want) + 1e-4
    bad = np.any(np.abs(got - want) > tol, axis=1)
    print(f"{name}: {bad.sum()} of {len(bad)} rays outside half precision; max rel err "
          f"{np.max(np.abs(got - want) / (np.abs(want) + 1e-3)):.2e}", flush=True)
    assert bad.mean() < 2e-3


@pytest.mark.parametrize("k", [0, 60, 119])
def test_grid_sky_and_disk_match_the_oracle(tracer, k):
    v = c5_view(k, 96, 54)
    out, host = traced(tracer, v)
    rgba = tracer.shade([v], out, sky=bsa.SKY_GRID, disk_temperature=8000.0).float().cpu().numpy()
    r_in = bsa.isco_radius(v.spin, True)
    want = O.shade(host, r_in, v.disk_outer, disk_temperature=8000.0, sky=1)
    assert np.all(rgba[:, 3] == 1.0)
    compare(rgba[:, :3], want, f"frame {k}, grid sky")
    cls = host["cls"] & 7
    assert np.all(rgba[cls == bsa.CAPTURED, :3] == 0.0)
    # disk rays glow, except at the faded edges and where g T is so cool that the colour
    # is below the smallest half
    assert (rgba[cls == bsa.DISK, :3].sum(axis=1) > 0.0).mean() > 0.9


def test_image_sky_matches_the_oracle(tracer):
    v = bsa.View(width=64, height=48, spin=0.6, r_cam=30.0, theta_cam=1.2, half_width=12.0, half_height=9.0,
                 disk_enabled=0)
    out, host = traced(tracer, v)
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 4, (32, 64, 3)).astype(np.float32)
    rgba = tracer.shade([v], out, sky=bsa.SKY_IMAGE, sky_image=img, sky_brightness=0.5).float().cpu().numpy()
    want = O.shade(host, 1.0, 20.0, sky=2, sky_image=img, sky_brightness=0.5)
    compare(rgba[:, :3], want, "image sky")
    assert (host["cls"] & 7 == bsa.ESCAPED).mean() > 0.5


def test_shading_reads_fp32_outputs_too(tracer):
    """kt_shade_views on the FP32 output fields: the colours of the FP64 outputs rounded to
    float (the same function of values within a float ulp: within 2 half ulps)."""
    import torch

    v = bsa.View(width=64, height=48, spin=0.9, r_cam=30.0, theta_cam=1.3, half_width=14.0, half_height=10.5)
    d = tracer.trace_views([v])
    f = tracer.trace_views([v], out=tracer.alloc_outputs(v.rays, f32=True))
    a = tracer.shade([v], d)
    b = tracer.shade([v], f)
    torch.cuda.synchronize()
    a = a.float().cpu().numpy()
    b = b.float().cpu().numpy()
    assert np.allclose(a, b, rtol=2e-3, atol=1e-6)
