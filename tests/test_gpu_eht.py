"""EHT radiative transfer on the GPU (kt_eht_render, kt_eht_trace_rays, kt_eht_flux,
kt_eht_calibrate) against the oracle's ko_eht_* on the same inputs.

The geodesic is the FP64 DP5 of the main tracer with every operation correctly rounded, so
the capture flag, the minimum radius and the step count are bit-identical to the oracle's.
The transfer calls exp, pow, cbrt, log and expm1, whose device and host versions may
differ in the last bit, so the intensity is held to 1e-11 relative (a pixel sums a few
hundred such terms).  The reference's own image test (tests/test_kerr.cpp:2322-2356:
calibrated to M87*'s 0.5 Jy, Doppler-brightened approaching side, shadow area equal to the
critical curve's) runs on the GPU path, and the traced shadow's edge is held to the
critical curve as tests/test_kerr.cpp:2281-2304 does.
"""
import math
import os

import numpy as np
import pytest
import torch

from black_simulation_amd import eht
from black_simulation_amd.api import TraceOpts
from oracle import oracle as O

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "eht_vectors.npz"))
M87 = eht.TARGETS[0]


def views():
    out = []
    for (a, inc, hw, nu, mass), flow in zip(G["trace_views"], G["trace_flows"]):
        out.append((eht.EhtView(a, inc, hw, 24, nu, mass), eht.Flow(*flow.tolist())))
    out.append((eht.EhtView(0.998, math.radians(80.0), 12.0, 24, 230e9, 6.5e9), eht.Flow(n0=3e5, height=0.4)))
    return out


def oracle_image(v, f):
    ov = O.eht_view(v.a, v.inclination, v.half_width, v.size, v.frequency, v.mass_solar)
    of = O.flow(f.n0, f.T0, f.beta, f.keplerian, f.height, f.sin_theta_b)
    o = O.make_opts()
    px = v.pixel()
    rays = []
    for row in range(v.size):
        for col in range(v.size):
            r = O.eht_trace(ov, of, o, -v.half_width + (col + 0.5) * px, v.half_width - (row + 0.5) * px)
            rays.append((r.intensity, r.captured, r.min_radius, r.steps))
    return np.array(rays)


def host(img):
    torch.cuda.synchronize()
    return {k: getattr(img, k).cpu().numpy().ravel() for k in ("intensity", "shadow", "min_radius", "steps")}


def assert_parity(got, want, what):
    assert np.array_equal(got["shadow"], want[:, 1].astype(np.uint8)), what
    assert np.array_equal(got["steps"], want[:, 3].astype(np.int32)), what
    assert np.array_equal(got["min_radius"], want[:, 2]), what
    scale = max(want[:, 0].max(), 1e-300)
    err = np.abs(got["intensity"] - want[:, 0])
    print(f"{what}: max rel err {np.max(err / np.maximum(np.abs(want[:, 0]), 1e-30 * scale)):.2e}, "
          f"lit {np.mean(want[:, 0] > 0):.2f}, captured {np.mean(want[:, 1]):.2f}", flush=True)
    assert np.all(err <= 1e-11 * np.abs(want[:, 0]) + 1e-14 * scale), what


@pytest.mark.parametrize("k", [0, 1, 2])
def test_render_matches_the_oracle(k):
    v, f = views()[k]
    img = eht.render(v, f, details=True)
    assert_parity(host(img), oracle_image(v, f), f"view {k}")


def test_explicit_rays_and_host_outputs_match():
    v, f = views()[0]
    ab = G["trace"][0][:, :2]
    dev = eht.trace_rays(v, f, ab)
    torch.cuda.synchronize()
    ov = O.eht_view(v.a, v.inclination, v.half_width, 8, v.frequency, v.mass_solar)
    of = O.flow(f.n0, f.T0, f.beta, f.keplerian, f.height, f.sin_theta_b)
    want = []
    for al, b in ab:
        r = O.eht_trace(ov, of, O.make_opts(), al, b)
        want.append((r.intensity, r.captured, r.min_radius, r.steps))
    got = {k: dev[k].cpu().numpy() for k in ("intensity", "shadow", "min_radius", "steps")}
    assert_parity(got, np.array(want), "explicit rays")
    one = eht.trace_ray(v, f, float(ab[5, 0]), float(ab[5, 1]))
    assert one["intensity"] == got["intensity"][5] and one["steps"] == got["steps"][5]
    assert one["captured"] == bool(got["shadow"][5])


def test_flux_is_a_fixed_order_sum():
    v, f = views()[1]
    img = eht.render(v, f)
    theta_g = eht.angular_scale(4.3e6, 8.28e3)
    j1 = eht.flux(img, theta_g)
    j2 = eht.flux(img, theta_g)
    assert j1 == j2
    inten = img.intensity.cpu().numpy()
    assert j1 == pytest.approx(O.eht_flux(inten, v.half_width, theta_g), rel=1e-13)


def test_calibrate_matches_the_oracle_and_the_reference():
    ci = G["calibrate_in"]
    v = eht.EhtView(ci[0], ci[1], ci[2], 48, ci[3], ci[4])
    f = eht.Flow(*G["calibrate_flow"].tolist())
    n0 = eht.calibrate(v, f, ci[5], ci[6])
    ov = O.eht_view(ci[0], ci[1], ci[2], 48, ci[3], ci[4])
    want = O.eht_calibrate(ov, O.flow(*G["calibrate_flow"].tolist()), O.make_opts(), ci[5], ci[6])
    print(f"n0: GPU {n0:.6e}, oracle {want:.6e}, reference {float(G['calibrate_n0']):.6e}", flush=True)
    assert n0 == pytest.approx(want, rel=1e-9)
    assert n0 == pytest.approx(float(G["calibrate_n0"]), rel=1e-2)


def test_m87_image_as_the_reference_tests_it():
    """tests/test_kerr.cpp:2322-2356: a = 0.9 at 60 degrees, calibrated on a 40-pixel view,
    rendered at 48: within 8% of 0.5 Jy, the approaching side 1.5 times the brighter, the
    shadow's area the critical curve's to 5%."""
    theta_g = eht.angular_scale(M87.mass_solar, M87.distance_pc)
    assert theta_g == pytest.approx(3.82, abs=0.01)
    v = eht.EhtView(a=0.9, inclination=math.radians(60.0), size=40, mass_solar=M87.mass_solar)
    f = eht.Flow()
    f.n0 = eht.calibrate(v, f, M87.flux, theta_g)
    v.size = 48
    img = eht.render(v, f)
    jy = eht.flux(img, theta_g)
    assert jy / M87.flux == pytest.approx(1.0, abs=0.08)
    inten = img.intensity.cpu().numpy()
    left, right = inten[:, :24].sum(), inten[:, 24:].sum()
    assert left > 1.5 * right
    al, be = O.critical_curve(0.9, math.radians(60.0), 400)
    area = 0.5 * abs(np.sum(al * np.roll(be, -1) - np.roll(al, -1) * be))
    shadow = img.shadow.cpu().numpy()
    assert shadow.sum() * v.pixel() ** 2 / area == pytest.approx(1.0, abs=0.05)


def test_traced_shadow_edge_follows_the_critical_curve():
    """tests/test_kerr.cpp:2281-2304 with the light on: along 8 directions per view, 4096
    radii in [1, 12] M; the last captured radius sits on the critical curve to 2e-3."""
    worst = 0.0
    radii = np.linspace(1.0, 12.0, 4096)
    for a, deg in ((0.9, 17.0), (0.9, 60.0), (0.0, 40.0)):
        v = eht.EhtView(a=a, inclination=math.radians(deg), size=8)
        al, be = O.critical_curve(a, v.inclination, 400)
        psis = [2 * math.pi * (k + 0.25) / 8 for k in range(8)]
        ab = np.concatenate([np.stack([radii * math.cos(p), radii * math.sin(p)], 1) for p in psis])
        out = eht.trace_rays(v, eht.Flow(n0=1e5), ab)
        cap = out["shadow"].cpu().numpy().reshape(8, -1)
        for k, p in enumerate(psis):
            edge = radii[np.nonzero(cap[k])[0].max()]
            want = O.critical_radius(al, be, p)
            worst = max(worst, abs(edge / want - 1.0))
    print(f"worst edge error {worst:.2e}", flush=True)
    assert worst < 2e-3


def test_edge_on_image_is_mirror_symmetric_at_full_size():
    """At 90 degrees the flow and the spacetime are symmetric under beta -> -beta, so a
    1024 x 1024 image is its own mirror image (the start's cos(pi/2) = 6e-17 and the
    step-size sequence differ between mirrored rays only by rounding)."""
    v = eht.EhtView(a=0.7, inclination=math.pi / 2, half_width=14.0, size=1024)
    img = eht.render(v, eht.Flow())
    inten = img.intensity.cpu().numpy()
    shadow = img.shadow.cpu().numpy()
    flip = inten[::-1]
    lit = inten > 1e-3 * inten.max()
    rel = np.abs(inten - flip)[lit] / inten[lit]
    print(f"mirror: median rel diff {np.median(rel):.2e}, p99.9 {np.quantile(rel, 0.999):.2e}, "
          f"shadow mismatch {np.mean(shadow != shadow[::-1]):.2e}", flush=True)
    assert np.median(rel) < 1e-9
    assert np.mean(rel > 1e-4) < 1e-3
    assert np.mean(shadow != shadow[::-1]) < 1e-3
    # not vacuous: the shadow and the lit flow both cover a good part of the image
    assert shadow.mean() > 0.05 and lit.mean() > 0.05, (shadow.mean(), lit.mean())


def test_non_finite_view_or_flow_is_refused():
    """an infinite frequency or a NaN flow parameter is KT_EINVAL, not an image of NaNs"""
    with pytest.raises(RuntimeError, match="must be finite"):
        eht.render(eht.EhtView(size=8, frequency=math.inf))
    with pytest.raises(RuntimeError, match="must be finite"):
        eht.render(eht.EhtView(size=8), eht.Flow(n0=math.nan))


def test_cancel_flag_stops_between_slices():
    """eht::render's and eht::calibrate's cancel (src/eht.cpp:330-331, 446): a flag clear
    throughout gives the uncancelled image bit for bit (slices trace the same rays); a flag
    set before the call traces nothing (all zero, cancelled) and calibrate returns its
    starting density; a flag set by another thread mid-render leaves whole slices of the
    image either exactly the full render's or zero, the zero ones after the traced ones."""
    import ctypes
    import threading
    import time

    v = eht.EhtView(0.9, math.radians(60.0), 12.0, 768, 230e9, 6.5e9)
    f = eht.Flow(n0=2e5)
    full = host(eht.render(v, f, details=True))
    flag = ctypes.c_int(0)
    img = eht.render(v, f, TraceOpts(cancel=flag), details=True)
    assert not img.cancelled
    clear = host(img)
    for k in full:
        assert np.array_equal(clear[k], full[k]), k
    flag.value = 1
    img = eht.render(v, f, TraceOpts(cancel=flag), details=True)
    assert img.cancelled
    for k, x in host(img).items():
        assert not x.any(), k
    assert eht.calibrate(eht.EhtView(0.9, 1.0, 16.0, 48), f, 0.5, 3.8, TraceOpts(cancel=flag)) == 2e5
    # from another thread, a few milliseconds in (a 768^2 render takes tens of ms)
    flag.value = 0
    setter = threading.Timer(0.004, lambda: setattr(flag, "value", 1))
    setter.start()
    t0 = time.perf_counter()
    img = eht.render(v, f, TraceOpts(cancel=flag), details=True)
    ms = 1e3 * (time.perf_counter() - t0)
    setter.join()
    got = host(img)
    n = v.size * v.size
    slice_ = max((n + 7) // 8, 1 << 17)
    done = []
    for s in range(0, n, slice_):
        part = {k: x[s:s + slice_] for k, x in got.items()}
        same = all(np.array_equal(part[k], full[k][s:s + slice_]) for k in part)
        zero = all(not x.any() for x in part.values())
        assert same or zero, s
        done.append(same and not zero)
    print(f"mid-render cancel after {ms:.1f} ms: cancelled {img.cancelled}, slices traced {done}", flush=True)
    assert done == sorted(done, reverse=True)  # traced slices first
    assert img.cancelled == (not all(done))
