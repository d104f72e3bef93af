"""kt_trace_views_noalloc, the entry point that allocates nothing, through the C ABI: its
outputs equal kt_trace_views' bit for bit on two calls in a row (the second smaller, so
the fixed scratch's counters and view table are reused and must start over), in FP64 and
in mixed precision; and what it refuses (more views than its fixed table holds, host
outputs, the surrogate)."""
import ctypes as C

import numpy as np
import pytest
import torch

import black_simulation_amd as bsa
from black_simulation_amd import _lib as K
from black_simulation_amd.configs import c1_view, c3_views

pytestmark = pytest.mark.gpu


def noalloc(tracer, views, opts, out=None):
    """one kt_trace_views_noalloc call on the current stream; returns (status, outputs)"""
    n = sum(v.rays for v in views)
    out = tracer.alloc_outputs(n) if out is None else out
    arr = (K.KtView * len(views))(*[v.to_c() for v in views])
    o = opts.to_c()
    s = tracer._out_struct(out)
    st = torch.cuda.current_stream(tracer.device)
    rc = K.lib().kt_trace_views_noalloc(tracer._ctx, arr, len(views), C.byref(o), C.byref(s),
                                        C.c_void_p(st.cuda_stream))
    return rc, out


def small(v, w, h):
    v.width, v.height = w, h
    return v


@pytest.mark.parametrize("precision", [bsa.PREC_FP64, bsa.PREC_MIXED])
def test_noalloc_equals_trace_views_on_two_calls(tracer, precision):
    big = c3_views(n=5, sizes=(24, 32, 40)) + [small(c1_view(), 48, 40)]
    little = [small(c1_view(), 20, 12)]
    opts = bsa.TraceOpts(precision=precision)
    for views in (big, little):
        n = sum(v.rays for v in views)
        print(f"noalloc vs trace_views: {len(views)} views, {n} rays, precision {precision}", flush=True)
        want = bsa.to_numpy(tracer.trace_views(views, opts))
        want_stats = tracer.stats()
        # outputs pre-filled with garbage: every field must be written
        out = tracer.alloc_outputs(n)
        for t in out.values():
            t.fill_(7)
        rc, out = noalloc(tracer, views, opts, out)
        assert rc == K.KT_OK, K.lib().kt_last_error_string(tracer._ctx)
        torch.cuda.synchronize()
        got = bsa.to_numpy(out)
        st = tracer.stats()
        for k in ("cls", "steps", "theta_inf", "phi_inf", "r_hit", "g"):
            assert np.array_equal(got[k], want[k], equal_nan=True), k
        assert st["rays"] == n and st["steps"] == want_stats["steps"] == int(want["steps"].sum())


def test_noalloc_refuses_what_it_cannot_hold(tracer):
    v = small(c1_view(), 8, 8)
    rc, _ = noalloc(tracer, [v] * 65, bsa.TraceOpts())
    assert rc == K.KT_EINVAL
    assert b"at most" in K.lib().kt_last_error_string(tracer._ctx)
    rc, _ = noalloc(tracer, [v], bsa.TraceOpts(use_surrogate=1))
    assert rc == K.KT_EINVAL
    # host outputs: the entry point writes device memory only
    arr = (K.KtView * 1)(v.to_c())
    o = bsa.TraceOpts().to_c()
    cls = np.zeros(v.rays, np.int8)
    s = K.KtRaysOut()
    s.cls = cls.ctypes.data
    s.on_device = 0
    rc = K.lib().kt_trace_views_noalloc(tracer._ctx, arr, 1, C.byref(o), C.byref(s), None)
    assert rc == K.KT_EINVAL
    # 64 views, the most it takes, still trace as kt_trace_views does
    views = [small(c1_view(), 4, 3) for _ in range(64)]
    rc, out = noalloc(tracer, views, bsa.TraceOpts())
    assert rc == K.KT_OK
    torch.cuda.synchronize()
    want = bsa.to_numpy(tracer.trace_views(views))
    assert np.array_equal(bsa.to_numpy(out)["cls"], want["cls"])
