"""The closed-form solution on the GPU (kt_trace_opts.closed_form = 1, kernel k_closed_form)
against the FP64 oracle (the DP5 integrator at rtol 1e-10 restating the reference's march).

Bars: every ray's class equal to the oracle's; the rays the closed form answers (flag
KT_FLAG_CLOSED_FORM, 0 steps) within 1e-6 relative in disk radius and g and 1e-6 rad in sky
direction, with medians below 1e-9 (the oracle is itself good to ~1e-9); the rays it
declines traced by the FP64 integrator, bit-equal to the oracle.  The FP32 pass
(closed_form = 2): within 1e-4 in disk radius and g, medians below 1e-6."""
import math

import numpy as np
import pytest

import black_simulation_amd as bsa
from black_simulation_amd.configs import C5_FRAMES, c1_view, c3_views, c5_view
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
          "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
          "escape_radius", "capture_factor")
CLOSED = bsa.TraceOpts(closed_form=1)


def oracle_view(v):
    return O.make_view(**{k: getattr(v, k) for k in FIELDS})


def sky_angle(t1, p1, t2, p2):
    v1 = np.stack([np.sin(t1) * np.cos(p1), np.sin(t1) * np.sin(p1), np.cos(t1)])
    v2 = np.stack([np.sin(t2) * np.cos(p2), np.sin(t2) * np.sin(p2), np.cos(t2)])
    return 2 * np.arcsin(np.minimum(1, np.linalg.norm(v1 - v2, axis=0) / 2))


def check(got, want, tag, min_answered=0.99):
    """one view's outputs against the oracle's render of it; returns the answered fraction"""
    cls = got["cls"] & 7
    closed = (got["cls"] & bsa.FLAG_CLOSED_FORM) != 0
    assert np.array_equal(cls, want["cls"]), f"{tag}: {int((cls != want['cls']).sum())} classes differ"
    assert (got["steps"][closed] == 0).all()
    fb = ~closed
    for k, w in (("steps", "steps"), ("r_hit", "r_hit"), ("g", "g")):
        assert np.array_equal(got[k][fb], want[w][fb], equal_nan=True), f"{tag}: declined rays' {k}"
    disk = closed & (cls == O.DISK)
    r_err = np.abs(got["r_hit"][disk] / want["r_hit"][disk] - 1)
    g_err = np.abs(got["g"][disk] / want["g"][disk] - 1)
    esc = closed & (cls == O.ESCAPED)
    d = sky_angle(got["theta_inf"][esc], got["phi_inf"][esc], want["theta"][esc], want["phi"][esc])
    q = lambda x: f"{np.median(x):.1e}/{x.max():.1e}" if len(x) else "-"  # noqa: E731
    print(f"{tag}: answered {closed.mean():.5f} of {len(cls)} (classes escaped/captured/disk "
          f"{np.bincount(cls, minlength=3)[:3].tolist()}); r_hit rel p50/max {q(r_err)}, g rel {q(g_err)}, "
          f"deflection rad {q(d)}", flush=True)
    for x in (r_err, g_err, d):
        if len(x):
            assert np.median(x) < 1e-9 and x.max() < 1e-6, tag
    assert closed.mean() >= min_answered, tag
    return closed.mean()


def test_c1_full_frame(tracer):
    """a = 0.9, r = 30 M, i = 80 deg, the whole 256x256 plane: every ray answered"""
    v = c1_view()
    print("closed form c1 256x256", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v], CLOSED))
    want = O.render(oracle_view(v), O.make_opts())
    assert check(got, want, "c1") == 1.0
    assert tracer.stats()["fallback_rays"] == 0


def test_orbit_frames_in_one_launch(tracer):
    """three frames of the a = 0.998 orbit (inclination 20, 53 and 85 deg) at 192x108 in one launch"""
    views = [c5_view(k, 192, 108) for k in (0, 60, C5_FRAMES - 1)]
    print("closed form: 3 orbit frames at 192x108", flush=True)
    got = bsa.to_numpy(tracer.trace_views(views, CLOSED))
    off = 0
    for i, v in enumerate(views):
        n = v.rays
        check({k: a[off:off + n] for k, a in got.items()}, O.render(oracle_view(v), O.make_opts()), f"orbit view {i}")
        off += n


def test_batch_of_64_views(tracer):
    """the 64 seeded views of the batch workload (spin, inclination, camera radius drawn) at
    1/16 of their size in one launch"""
    views = c3_views(sizes=(8, 16, 32, 64, 128))
    print("closed form: 64 batch views", flush=True)
    got = bsa.to_numpy(tracer.trace_views(views, CLOSED))
    off = 0
    for i, v in enumerate(views):
        n = v.rays
        check({k: a[off:off + n] for k, a in got.items()}, O.render(oracle_view(v), O.make_opts()), f"batch view {i}")
        off += n


@pytest.mark.parametrize("case", ["schwarzschild", "retrograde", "inner6", "nodisk", "edge_on", "near_axis",
                                  "far", "negative_spin"])
def test_edge_views(tracer, case):
    """spins, disks and cameras the orbit does not have: a = 0; a retrograde disk; a disk from
    r = 6 M; no disk; the camera in the equatorial plane (mu = 0 at launch); 0.06 deg off the
    axis; a camera at 1000 M; a hole spinning the other way"""
    base = dict(width=48, height=48, spin=0.9, r_cam=30.0, theta_cam=math.radians(70.0), half_width=10.0,
                half_height=10.0, disk_outer=20.0)
    base.update({"schwarzschild": dict(spin=0.0), "retrograde": dict(disk_prograde=0),
                 "inner6": dict(disk_inner=6.0), "nodisk": dict(disk_enabled=0),
                 "edge_on": dict(theta_cam=math.pi / 2, height=47), "near_axis": dict(theta_cam=1e-3),
                 "far": dict(r_cam=1000.0), "negative_spin": dict(spin=-0.7)}[case])
    v = bsa.View(**base)
    print(f"closed form: edge view {case}", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v], CLOSED))
    check(got, O.render(oracle_view(v), O.make_opts()), case, min_answered=0.9)


def test_pinhole_cameras(tracer):
    """the simulator's camera (static ZAMO frame): looking at the hole from 32 M, and from 8 M
    looking past it so that part of the field heads outward (those rays escape without a
    turning point; the few left to the integrator escape with R's second pair of roots
    complex)"""
    views = [bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=48, height=32, spin=0.94, r_cam=32.0, theta_cam=1.48,
                      yaw=0.05, pitch=-0.02, roll=0.1, fov_deg=60.0, disk_outer=20.0),
             bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=48, height=32, spin=0.7, r_cam=8.0, theta_cam=1.2,
                      yaw=1.2, pitch=0.3, fov_deg=120.0, disk_outer=20.0)]
    print("closed form: pinhole cameras", flush=True)
    got = bsa.to_numpy(tracer.trace_views(views, CLOSED))
    off = 0
    for i, v in enumerate(views):
        n = v.rays
        check({k: a[off:off + n] for k, a in got.items()}, O.render(oracle_view(v), O.make_opts()), f"pinhole {i}",
              min_answered=0.99)
        off += n


def check32(got, want, tag):
    """the FP32 pass: classes equal to the oracle's; disk radius and g within 1e-4 (median
    1e-6); escapes solved in FP64 (1e-6 rad); declined rays bit-equal to the oracle"""
    cls = got["cls"] & 7
    closed = (got["cls"] & bsa.FLAG_CLOSED_FORM) != 0
    assert np.array_equal(cls, want["cls"]), f"{tag}: {int((cls != want['cls']).sum())} classes differ"
    fb = ~closed
    for k in ("steps", "r_hit", "g"):
        assert np.array_equal(got[k][fb], want[k][fb], equal_nan=True), f"{tag}: declined rays' {k}"
    disk = closed & (cls == O.DISK)
    r_err = np.abs(got["r_hit"][disk] / want["r_hit"][disk] - 1)
    g_err = np.abs(got["g"][disk] / want["g"][disk] - 1)
    esc = closed & (cls == O.ESCAPED)
    d = sky_angle(got["theta_inf"][esc], got["phi_inf"][esc], want["theta"][esc], want["phi"][esc])
    q = lambda x: f"{np.median(x):.1e}/{np.quantile(x, 0.99):.1e}/{x.max():.1e}" if len(x) else "-"  # noqa: E731
    print(f"{tag} (FP32 pass): answered {closed.mean():.5f}; r_hit rel p50/p99/max {q(r_err)}, g rel {q(g_err)}, "
          f"deflection rad {q(d)}", flush=True)
    for x in (r_err, g_err):
        if len(x):
            assert np.median(x) < 1e-6 and x.max() < 1e-4, tag
    if len(d):
        assert d.max() < 1e-6, tag


def test_fp32_pass(tracer):
    """closed_form = 2: the c1 plane, three orbit frames and the 64 batch views"""
    FAST = bsa.TraceOpts(closed_form=2)
    v = c1_view()
    check32(bsa.to_numpy(tracer.trace_views([v], FAST)), O.render(oracle_view(v), O.make_opts()), "c1")
    for views, tag in (([c5_view(k, 192, 108) for k in (0, 60, C5_FRAMES - 1)], "orbit"),
                       (c3_views(sizes=(8, 16, 32, 64, 128)), "batch")):
        got = bsa.to_numpy(tracer.trace_views(views, FAST))
        off = 0
        for i, vv in enumerate(views):
            n = vv.rays
            check32({k: a[off:off + n] for k, a in got.items()}, O.render(oracle_view(vv), O.make_opts()),
                    f"{tag} view {i}")
            off += n


def test_explicit_rays(tracer):
    """kt_trace_rays with explicit (alpha, beta), as eht::traceRay"""
    v = c1_view()
    rng = np.random.default_rng(7)
    ab = rng.uniform(-12, 12, size=(300, 2))
    got = bsa.to_numpy(tracer.trace_params(v, ab, CLOSED))
    want = O.trace_params(oracle_view(v), O.make_opts(), ab)
    check(got, want, "explicit rays")


def test_plan_replays_equal_the_eager_call(tracer):
    """a captured plan replayed three times, eager work between replays: every replay equal
    to the eager closed-form trace bit for bit"""
    import torch

    views = [c5_view(k, 160, 90) for k in (10, 100)]
    eager = bsa.to_numpy(tracer.trace_views(views, CLOSED))
    out = tracer.alloc_outputs(sum(v.rays for v in views))
    plan = tracer.plan(views, CLOSED, out=out)
    try:
        for rep in range(3):
            for k in out:
                out[k].fill_(0)
            tracer.trace_views([c1_view()], CLOSED)  # eager work between replays
            plan.launch()
            torch.cuda.synchronize()
            got = bsa.to_numpy(out)
            for k in ("cls", "theta_inf", "phi_inf", "r_hit", "g", "steps"):
                assert np.array_equal(got[k], eager[k], equal_nan=True), (rep, k)
    finally:
        plan.close()


@pytest.mark.parametrize("closed", [1, 2])
def test_surrogate_with_closed_form_values(tracer, closed):
    """use_surrogate with closed_form: the network's escaped and disk answers carry the closed
    form's values, bit-equal to the closed-form path's at the same precision; every class as
    the oracle's on >= 99.9 % of rays (the network's captured answers are its own)"""
    from black_simulation_amd.train_surrogate import FAST_WEIGHTS, load_into

    load_into(tracer, FAST_WEIGHTS)
    views = [c5_view(k, 192, 108) for k in (5, 70, 115)]
    print("surrogate + closed-form values: 3 orbit frames", flush=True)
    sur = bsa.to_numpy(tracer.trace_views(views, bsa.TraceOpts(use_surrogate=1, closed_form=closed, rtol=1e-6,
                                                                atol=1e-7)))
    cf = bsa.to_numpy(tracer.trace_views(views, bsa.TraceOpts(closed_form=closed)))
    cls = sur["cls"] & 7
    polished = (sur["cls"] & bsa.FLAG_CLOSED_FORM) != 0
    answered = (sur["cls"] & bsa.FLAG_SURROGATE) != 0
    assert (polished <= answered).all()
    assert np.isin(cls[polished], [O.ESCAPED, O.DISK]).all()
    assert ((cls[answered & ~polished]) == O.CAPTURED).all()
    for k in ("theta_inf", "phi_inf", "r_hit", "g"):
        assert np.array_equal(sur[k][polished], cf[k][polished], equal_nan=True), k
    assert np.array_equal(cls[polished], cf["cls"][polished] & 7)
    off, agree, n = 0, 0, 0
    for v in views:
        want = O.render(oracle_view(v), O.make_opts())
        agree += int((cls[off:off + v.rays] == want["cls"]).sum())
        n += v.rays
        off += v.rays
    print(f"surrogate + closed form: answered {answered.mean():.4f}, closed-form values on {polished.mean():.4f}, "
          f"class agreement {agree / n:.6f}", flush=True)
    assert polished.mean() > 0.5 and agree / n >= 0.999


def test_close_cameras(tracer):
    """Bardeen planes seen from inside the photon region (r = 1.8-3.2 M): some rays start
    between the roots r2 and r3 of R (the ray falls to r2 below the capture radius, after
    turning at r3 if it starts outward), others beyond r4; pixels whose launch R(r_cam) < 0
    names no photon are left to the integrator"""
    views = [bsa.View(width=40, height=40, spin=0.9, r_cam=2.6, theta_cam=1.2, half_width=7.0, half_height=7.0),
             bsa.View(width=40, height=40, spin=0.5, r_cam=3.2, theta_cam=0.7, half_width=8.0, half_height=8.0,
                      disk_inner=2.0),
             bsa.View(width=40, height=40, spin=0.998, r_cam=1.8, theta_cam=1.4, half_width=6.0, half_height=6.0),
             bsa.View(width=40, height=40, spin=0.9, r_cam=2.2, theta_cam=1.5, half_width=5.0, half_height=5.0)]
    for i, v in enumerate(views):
        print(f"close camera {i}", flush=True)
        want = O.render(oracle_view(v), O.make_opts())
        check(bsa.to_numpy(tracer.trace_views([v], CLOSED)), want, f"close camera {i}", min_answered=0.3)
        check32(bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(closed_form=2))), want, f"close camera {i}")


def test_disk_near_the_photon_orbit(tracer):
    """a counter-rotating disk whose inner edge lies below its ISCO (a negative spin with the
    prograde disk and an inner edge at 2.92 M, or the ISCO formula's prograde radius): its
    emitters reach the photon orbit, where g falls to zero as steeply as sqrt(d), d =
    r^1.5 - 3 r^0.5 + 2 s a, and whether a circular orbit exists flips with the last digits of
    the crossing radius.  The FP32 pass hands those rays to FP64: every class equals the
    oracle's, and g stays within the FP32 bar"""
    views = [bsa.View(width=48, height=48, spin=-0.3224500706761697, r_cam=397.591112698606,
                      theta_cam=1.1732759034435825, phi_cam=0.4978852283607192, half_width=8.349956684111284,
                      half_height=8.349956684111284, disk_outer=31.167540346017052, disk_inner=2.9200540150299337),
             bsa.View(width=48, height=48, spin=-0.7967212696420716, r_cam=667.9740974565108,
                      theta_cam=0.2299825628108736, phi_cam=1.2866744128204795, half_width=12.009798480309898,
                      half_height=12.009798480309898, disk_outer=15.840551924666542)]
    for i, v in enumerate(views):
        print(f"disk near the photon orbit {i}", flush=True)
        want = O.render(oracle_view(v), O.make_opts())
        check32(bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(closed_form=2))), want, f"near photon orbit {i}")
        check(bsa.to_numpy(tracer.trace_views([v], CLOSED)), want, f"near photon orbit {i} FP64", min_answered=0.9)


@pytest.mark.parametrize("closed", [1, 2])
def test_observer_cameras(tracer, closed):
    """the renderer's camera carried by any observer (KT_CAMERA_OBSERVER, phi_cam the ingoing
    Kerr-Schild phi): a camera falling from rest at infinity at r = 10 M and a boosted ZAMO
    at 15 M are answered by the closed form; one falling inside the horizon (r = 1.2 M) is
    left whole to the integrator, bit-equal to the oracle"""
    a = 0.9
    views = [bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=40, height=30, spin=a, r_cam=10.0, theta_cam=1.0,
                      phi_cam=0.4, fov_deg=70.0, yaw=0.1, disk_outer=20.0,
                      u_obs=bsa.geodesic_four_velocity(a, 1.0, 0.0, 0.0, 10.0, 1.0, -1, 1)),
             bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=40, height=30, spin=a, r_cam=15.0, theta_cam=1.3,
                      phi_cam=-0.7, fov_deg=50.0, pitch=-0.05, disk_outer=20.0, drift=(0.3, 0.0, 0.5)),
             bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=24, height=16, spin=a, r_cam=1.2, theta_cam=1.05,
                      fov_deg=120.0, disk_outer=20.0,
                      u_obs=bsa.geodesic_four_velocity(a, 1.0, 0.0, 0.0, 1.2, 1.05, -1, 1))]
    fields = FIELDS + ("u_obs", "drift", "frame")
    for i, v in enumerate(views):
        print(f"closed form ({closed}): observer camera {i}", flush=True)
        got = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(closed_form=closed)))
        want = O.render(O.make_view(**{k: getattr(v, k) for k in fields}), O.make_opts())
        if closed == 1:
            answered = check(got, want, f"observer {i}", min_answered=0.0)
        else:
            check32(got, want, f"observer {i}")
            answered = ((got["cls"] & bsa.FLAG_CLOSED_FORM) != 0).mean()
        assert answered > 0.9 if i < 2 else answered == 0.0, (i, answered)


def test_closed_form_refusals(tracer):
    """what the closed form does not take is refused with KT_EINVAL, not silently traced"""
    with pytest.raises(RuntimeError, match="closed_form must be 0, 1 or 2"):
        tracer.trace_views([c1_view()], bsa.TraceOpts(closed_form=3))
    with pytest.raises(RuntimeError, match="report"):
        tracer.trace_params(c1_view(), [[1.0, 2.0]], CLOSED, report=True)
