"""The workloads the tracer is measured and checked on, by name and size.

    c1_view()       a=0.9, Bardeen camera r=30 M, i=80 deg, 256x256 over 24x24 M,
                    disk r_ISCO..20 M (the CPU-runnable case)
    c2_view()       a=0.998, r=50 M, i=85 deg, 1920x1080 over a 30 M wide plane,
                    disk r_ISCO..40 M (the single large frame)
    c3_views(seed)  64 seeded views of uneven size and spin in one batch
    c5_view(k, res) frame k of the 120-frame a=0.998 orbit at r=40 M
"""
from __future__ import annotations

import math

import numpy as np

from .api import View, shadow_radius

C3_SEED = 0x5EED0003


def c1_view() -> View:
    return View(width=256, height=256, spin=0.9, r_cam=30.0, theta_cam=math.radians(80.0), phi_cam=0.0,
                half_width=12.0, half_height=12.0, disk_inner=0.0, disk_outer=20.0, escape_radius=1000.0,
                capture_factor=1.01)


def c2_view(width: int = 1920, height: int = 1080) -> View:
    hw = 15.0
    hh = hw * height / width
    return View(width=width, height=height, spin=0.998, r_cam=50.0, theta_cam=math.radians(85.0), phi_cam=0.0,
                half_width=hw, half_height=hh, disk_inner=0.0, disk_outer=40.0, escape_radius=1000.0,
                capture_factor=1.01)


def c3_views(seed: int = C3_SEED, n: int = 64, sizes=(128, 256, 512, 1024, 2048)) -> list[View]:
    """Spin U[0, 0.998], inclination U[5, 85] deg, r_cam = 10^U[1, 3], square size from
    `sizes`, half-width 1.5x the mean critical radius.  Drawn with numpy's MT19937."""
    rng = np.random.Generator(np.random.MT19937(seed))
    views = []
    for _ in range(n):
        a = float(rng.uniform(0.0, 0.998))
        inc = math.radians(float(rng.uniform(5.0, 85.0)))
        r_cam = float(10.0 ** rng.uniform(1.0, 3.0))
        size = int(sizes[int(rng.integers(0, len(sizes)))])
        hw = 1.5 * shadow_radius(max(a, 1e-7), inc, samples=200, directions=72)
        views.append(View(width=size, height=size, spin=a, r_cam=r_cam, theta_cam=inc, half_width=hw,
                          half_height=hw, disk_inner=0.0, disk_outer=20.0, escape_radius=1000.0,
                          capture_factor=1.01))
    return views


C5_FRAMES = 120


def c5_view(k: int, width: int = 1920, height: int = 1080) -> View:
    """Frame k of the orbit: phi_k = 2 pi k / 120, inclination 20 + 65 k / 119 deg.  The
    half-width is 1.5x the frame's mean critical radius (the batch-view rule), so every
    pixel lies inside the range the surrogate is trained on."""
    inc = math.radians(20.0 + 65.0 * k / (C5_FRAMES - 1))
    hw = 1.5 * shadow_radius(0.998, inc, samples=200, directions=72)
    return View(width=width, height=height, spin=0.998, r_cam=40.0,
                theta_cam=inc, phi_cam=2.0 * math.pi * k / C5_FRAMES,
                half_width=hw, half_height=hw * height / width, disk_inner=0.0, disk_outer=20.0,
                escape_radius=1000.0, capture_factor=1.01)


def infall_view(width: int = 1920, height: int = 1080, r: float = 1.2, spin: float = 0.9) -> View:
    """A camera falling from rest at infinity (E = 1, L = Q = 0), here inside the horizon
    of a = 0.9 (r+ = 1.436 M) at r = 1.2 M, 60 degrees off the axis, looking down at the
    hole with a 120-degree field (disk r_ISCO..20 M, escape 1000 M)."""
    from .api import geodesic_four_velocity
    from ._lib import CAMERA_OBSERVER

    th = math.radians(60.0)
    u = geodesic_four_velocity(spin, 1.0, 0.0, 0.0, r, th, -1, 1)
    return View(camera_kind=CAMERA_OBSERVER, width=width, height=height, spin=spin, r_cam=r, theta_cam=th,
                phi_cam=0.0, fov_deg=120.0, disk_inner=0.0, disk_outer=20.0, escape_radius=1000.0,
                capture_factor=1.01, u_obs=u)
