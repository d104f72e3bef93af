"""Naked singularities (1 < a <= 1.5) on the GPU: the FP64 integrator against the oracle
bit for bit (Bardeen and pinhole views, rays through the ring's disk to r < 0 among them),
the mixed path against it, the stand-ins handing every such ray to the integrator, and what
is refused.  The oracle's naked branches are pinned to the reference in
tests/test_oracle_naked.py."""
import ctypes as C
import math

import numpy as np
import pytest

import black_simulation_amd as bsa
from black_simulation_amd import _lib as K
from oracle import oracle as O
from test_gpu_parity import compare, oracle_view

pytestmark = pytest.mark.gpu


def naked_views():
    """a Bardeen view wide enough to hold the rays that reach r = 0 (Q < 0) and the disk down
    to the ISCO inside r = M, and a pinhole camera looking at the ring from near the plane"""
    return [bsa.View(width=48, height=48, spin=1.2, r_cam=30.0, theta_cam=math.radians(60.0), half_width=3.0,
                     half_height=3.0, disk_outer=20.0),
            bsa.View(width=40, height=30, spin=1.5, r_cam=25.0, theta_cam=math.radians(80.0), half_width=14.0,
                     half_height=10.0),
            bsa.View(camera_kind=bsa.CAMERA_PINHOLE, width=40, height=40, spin=1.05, r_cam=15.0,
                     theta_cam=math.radians(75.0), fov_deg=30.0, disk_outer=20.0)]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_fp64_bit_exact(tracer, k):
    v = naked_views()[k]
    print(f"naked view {k}: a={v.spin} kind={v.camera_kind}", flush=True)
    got = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts()))
    want = O.render(oracle_view(v), O.make_opts())
    compare(got, want, f"naked-{k}")
    if k == 0:  # rays through the ring's disk are captured, the disk is hit inside r = M
        assert (want["cls"] == O.CAPTURED).sum() > 20
        assert np.nanmin(want["r_hit"]) < 1.0


def test_mixed_against_fp64(tracer):
    views = naked_views()
    ref = bsa.to_numpy(tracer.trace_views(views, bsa.TraceOpts()))
    got = bsa.to_numpy(tracer.trace_views(views, bsa.TraceOpts(precision=bsa.PREC_MIXED)))
    agree = np.mean((got["cls"] & 7) == (ref["cls"] & 7))
    disk = ((got["cls"] & 7) == O.DISK) & ((ref["cls"] & 7) == O.DISK)
    g_err = np.abs(got["g"][disk] / ref["g"][disk] - 1)
    print(f"naked mixed: class agreement {agree:.5f}, g rel max {g_err.max():.2e}", flush=True)
    assert agree >= 0.999 and g_err.max() < 1e-4


def test_stand_ins_hand_naked_rays_to_the_integrator(tracer):
    """the closed form declines every ray past a = 1 and the surrogate's domain ends at 0.998:
    the integrator traces them all, bit-equal to the plain FP64 call"""
    from black_simulation_amd.train_surrogate import FAST_WEIGHTS, load_into

    v = naked_views()[0]
    ref = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts()))
    cf = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(closed_form=2)))
    assert not (cf["cls"] & bsa.FLAG_CLOSED_FORM).any()
    for key in ("cls", "steps", "r_hit", "g"):
        assert np.array_equal(cf[key], ref[key], equal_nan=True), key
    load_into(tracer, FAST_WEIGHTS)
    sur = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(use_surrogate=1)))
    assert not (sur["cls"] & bsa.FLAG_SURROGATE).any()
    assert np.array_equal(sur["cls"] & 7, ref["cls"] & 7)


def test_refusals(tracer):
    lib = K.lib()
    out = K.KtRaysOut()
    o = bsa.TraceOpts().to_c()
    for v in (bsa.View(spin=1.6), bsa.View(spin=-1.2),
              bsa.View(camera_kind=bsa.CAMERA_OBSERVER, width=4, height=4, spin=1.2, r_cam=10.0, theta_cam=1.2,
                       fov_deg=40.0)):
        assert lib.kt_trace_views(tracer._ctx, C.byref(v.to_c()), 1, C.byref(o), C.byref(out), None) == K.KT_EINVAL
        print(K.lib().kt_last_error_string(tracer._ctx).decode(), flush=True)
