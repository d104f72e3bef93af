"""Frame driver for a camera orbit, and per-ray error statistics against a reference render.

The frame loop of the simulator (Renderer::render, src/renderer.cpp:336-379) draws one
frame per camera pose.  Here every pose of an orbit is prepared once as a kt_plan (view
table uploaded, launch sequence captured in a hipGraph) and a frame is one graph replay.
With use_surrogate set the graph holds the network, the gate and the integrator over the
rays the gate rejected.  Consecutive frames go to `streams` HIP streams in turn, each
with its own outputs, so one frame's slowest rays overlap the next frame's bulk work.
"""
from __future__ import annotations

import math
from typing import Sequence

from . import _lib as K
from .api import TraceOpts, Tracer, View


class Orbit:
    """One plan per frame; a frame is a view or a list of views traced in one launch (all
    frames of equal ray count).  Frame k writes outputs[k % streams] on streams[k % streams]."""

    def __init__(self, tracer: Tracer, frames: Sequence, opts: TraceOpts, streams: int = 1):
        frames = [list(f) if isinstance(f, (list, tuple)) else [f] for f in frames]
        n = sum(v.rays for v in frames[0])
        if any(sum(v.rays for v in f) != n for f in frames):
            raise ValueError("orbit frames must have equal ray counts")
        torch = tracer.torch
        self.tracer, self.frames, self.rays = tracer, frames, n
        self.streams = [torch.cuda.current_stream(tracer.device)] + \
                       [torch.cuda.Stream(tracer.device) for _ in range(streams - 1)]
        self.outputs = [tracer.alloc_outputs(n) for _ in range(streams)]
        self.plans = [tracer.plan(f, opts, out=self.outputs[k % streams]) for k, f in enumerate(frames)]

    @property
    def out(self) -> dict:
        return self.outputs[0]

    def render(self, k: int) -> dict:
        """Frame k, asynchronously on its stream; returns its outputs (valid once that
        stream has reached this point, until frame k + streams)."""
        s = k % len(self.streams)
        return self.plans[k].launch(self.streams[s])

    def stats(self, k: int) -> dict:
        """kt_plan_stats of frame k's last replay (waits for it)."""
        return self.plans[k].stats()

    def synchronize(self):
        for s in self.streams:
            s.synchronize()

    def close(self):
        for p in self.plans:
            p.close()
        self.plans = []


def _sky(theta, phi):
    st = theta.sin()
    return st * phi.cos(), st * phi.sin(), theta.cos()


def errors(out: dict, ref: dict, flags: int = K.FLAG_SURROGATE) -> dict:
    """Per-ray comparison of `out` with `ref` (device tensors of one frame).

    Returns device tensors: `agree` (class match, bool per ray), `answered` (ray carries one
    of `flags`: by default the surrogate's; the closed form's for its path), and for rays
    whose class matches: `defl` (angle between the two sky directions, rad) and
    `defl_answered` for escaped rays, `g_rel` / `r_rel` (relative errors of g and r_hit) and
    `*_answered` for disk rays.  `counts` (host ints) holds, per class of the reference, its
    rays, how many of them were answered and how many of those got the reference's class."""
    c = out["cls"].to(dtype=ref["cls"].dtype) & 7
    cr = ref["cls"] & 7
    ans = (out["cls"].to(dtype=ref["cls"].dtype) & flags) != 0
    agree = c == cr
    esc = agree & (cr == K.ESCAPED)
    x1, y1, z1 = _sky(out["theta_inf"][esc], out["phi_inf"][esc])
    x2, y2, z2 = _sky(ref["theta_inf"][esc], ref["phi_inf"][esc])
    cx, cy, cz = y1 * z2 - z1 * y2, z1 * x2 - x1 * z2, x1 * y2 - y1 * x2
    cross = (cx * cx + cy * cy + cz * cz).sqrt()
    dot = x1 * x2 + y1 * y2 + z1 * z2
    defl = cross.atan2(dot)
    disk = agree & (cr == K.DISK)
    g_rel = (out["g"][disk] / ref["g"][disk] - 1.0).abs()
    r_rel = (out["r_hit"][disk] / ref["r_hit"][disk] - 1.0).abs()
    counts = {}
    for name, k in (("escaped", K.ESCAPED), ("captured", K.CAPTURED), ("disk", K.DISK), ("unfinished", K.UNFINISHED)):
        of = cr == k
        counts[name] = {"rays": int(of.sum()), "answered": int((of & ans).sum()),
                        "answered_agree": int((of & ans & agree).sum())}
    return {"agree": agree, "answered": ans, "defl": defl, "defl_answered": defl[ans[esc]],
            "g_rel": g_rel, "g_rel_answered": g_rel[ans[disk]], "r_rel": r_rel, "r_rel_answered": r_rel[ans[disk]],
            "counts": counts}


def quantiles(x, qs=(0.5, 0.99, 0.999)) -> list:
    """Quantiles of a 1-D device tensor by sorting (any length); NaN when empty."""
    n = x.numel()
    if n == 0:
        return [math.nan for _ in qs]
    s, _ = x.sort()
    return [float(s[min(n - 1, int(q * (n - 1) + 0.5))]) for q in qs]
