"""What the Event Horizon Telescope sees, on the GPU: the reference's eht namespace
(src/eht.h:100-153) over libkerr_trace.so.

    render(view, flow)                 ~ eht::render      (src/eht.cpp:321-347)
    trace_ray(view, flow, alpha, beta) ~ eht::traceRay   (src/eht.cpp:260-319, with the light)
    trace_rays(view, flow, ab)         ~ many traceRay calls in one launch
    flux(image, theta_g)               ~ eht::flux        (src/eht.cpp:349-354)
    calibrate(view, flow, jy, theta_g) ~ eht::calibrate   (src/eht.cpp:432-457)
    angular_scale(mass, distance)      ~ eht::angularScale (src/eht.cpp:46-48)
    TARGETS                            ~ eht::targets     (src/eht.cpp:33-44)

Every pixel is traced in FP64 by the DP5 engine and the flow's thermal synchrotron light
gathered after each accepted step inside 60 M.  Images stay on the device (torch tensors)
unless asked for on the host.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib as K
from .api import TraceOpts, Tracer, _check


@dataclasses.dataclass
class Target:
    name: str
    mass_solar: float
    distance_pc: float
    inclination_deg: float  # of the spin axis to the line of sight
    flux: float             # Jy at 230 GHz
    ring_uas: float
    ring_error_uas: float


# the two holes the EHT imaged: M87* (EHT Collaboration 2019, ApJL 875, L1 and L6; the spin
# axis points away from us, so 180 - 17 degrees) and Sgr A* (EHT Collaboration 2022, ApJL
# 930, L12; fairly face-on)
TARGETS = (
    Target("M87*", 6.5e9, 16.8e6, 163.0, 0.5, 42.0, 3.0),
    Target("Sgr A*", 4.3e6, 8.28e3, 30.0, 2.4, 51.8, 2.3),
)


@dataclasses.dataclass
class Flow:
    """eht::Flow (src/eht.h:104-111): a hot, thick flow in cgs."""

    n0: float = 1e5
    T0: float = 4e11
    beta: float = 10.0
    keplerian: float = 0.9
    height: float = 1.0
    sin_theta_b: float = 0.8165

    def to_c(self) -> K.KtEhtFlow:
        return K.KtEhtFlow(self.n0, self.T0, self.beta, self.keplerian, self.height, self.sin_theta_b)


@dataclasses.dataclass
class EhtView:
    """eht::View (src/eht.h:118-125)."""

    a: float = 0.9
    inclination: float = 0.3
    half_width: float = 16.0
    size: int = 128
    frequency: float = 230e9
    mass_solar: float = 6.5e9

    def to_c(self) -> K.KtEhtView:
        return K.KtEhtView(self.a, self.inclination, self.half_width, self.size, self.frequency, self.mass_solar)

    def pixel(self) -> float:
        return 2.0 * self.half_width / self.size


@dataclasses.dataclass
class Image:
    """eht::Image (src/eht.h:132-138): rows from the top (beta = +half_width)."""

    size: int
    half_width: float
    intensity: object  # (size, size) float64, torch tensor on the device or numpy
    shadow: object     # (size, size) uint8
    min_radius: object = None
    steps: object = None
    cancelled: bool = False  # the opts' cancel flag stopped the render: rows not traced are zero

    def pixel(self) -> float:
        return 2.0 * self.half_width / self.size


def angular_scale(mass_solar: float, distance_pc: float) -> float:
    """theta_g = G M / (c^2 D) in microarcseconds."""
    return K.lib().kt_eht_angular_scale(mass_solar, distance_pc)


_default = None


def _tracer(tracer: Tracer | None) -> Tracer:
    global _default
    if tracer is not None:
        return tracer
    if _default is None:
        _default = Tracer()
    return _default


def _opts(opts: TraceOpts | None):
    return (opts or TraceOpts()).to_c()


def _stream(t: Tracer, stream):
    st = stream if stream is not None else t.torch.cuda.current_stream(t.device)
    return C.c_void_p(st.cuda_stream)


def _alloc(t: Tracer, n: int, details: bool):
    torch = t.torch
    dev = f"cuda:{t.device}"
    out = {"intensity": torch.empty(n, dtype=torch.float64, device=dev),
           "shadow": torch.empty(n, dtype=torch.uint8, device=dev)}
    if details:
        out["min_radius"] = torch.empty(n, dtype=torch.float64, device=dev)
        out["steps"] = torch.empty(n, dtype=torch.int32, device=dev)
    s = K.KtEhtOut(out["intensity"].data_ptr(), out["shadow"].data_ptr(),
                   out["min_radius"].data_ptr() if details else None, out["steps"].data_ptr() if details else None, 1)
    return out, s


def render(view: EhtView, flow: Flow | None = None, opts: TraceOpts | None = None, details: bool = False,
           tracer: Tracer | None = None, stream=None) -> Image:
    """kt_eht_render: the view's size^2 pixels on the device (asynchronous on `stream`)."""
    t = _tracer(tracer)
    out, s = _alloc(t, view.size * view.size, details)
    f = (flow or Flow()).to_c()
    v, o = view.to_c(), _opts(opts)
    rc = K.lib().kt_eht_render(t._ctx, C.byref(v), C.byref(f), C.byref(o), C.byref(s), _stream(t, stream))
    if rc != K.KT_ECANCELLED:
        _check(t._ctx, rc, "kt_eht_render")
    sq = {k: x.view(view.size, view.size) for k, x in out.items()}
    return Image(view.size, view.half_width, sq["intensity"], sq["shadow"], sq.get("min_radius"), sq.get("steps"),
                 cancelled=rc == K.KT_ECANCELLED)


def trace_rays(view: EhtView, flow: Flow | None, ab, opts: TraceOpts | None = None, tracer: Tracer | None = None,
               stream=None) -> dict:
    """kt_eht_trace_rays: eht::traceRay with the light for n (alpha, beta) pairs (in M), one
    launch; returns device tensors intensity, shadow (captured), min_radius, steps."""
    t = _tracer(tracer)
    torch = t.torch
    ab_t = torch.as_tensor(np.ascontiguousarray(ab, np.float64).reshape(-1, 2)).to(f"cuda:{t.device}")
    n = ab_t.shape[0]
    out, s = _alloc(t, n, True)
    f = (flow or Flow()).to_c()
    v, o = view.to_c(), _opts(opts)
    _check(t._ctx, K.lib().kt_eht_trace_rays(t._ctx, C.byref(v), C.byref(f), C.c_void_p(ab_t.data_ptr()), n,
                                             C.byref(o), C.byref(s), _stream(t, stream)), "kt_eht_trace_rays")
    out["_ab"] = ab_t  # keeps the input alive until the caller synchronises
    return out


def trace_ray(view: EhtView, flow: Flow | None, alpha: float, beta: float, opts: TraceOpts | None = None,
              tracer: Tracer | None = None) -> dict:
    """eht::Ray of one pixel: intensity, captured, min_radius, steps (host values)."""
    t = _tracer(tracer)
    intensity = np.zeros(1)
    shadow = np.zeros(1, np.uint8)
    rmin = np.zeros(1)
    steps = np.zeros(1, np.int32)
    s = K.KtEhtOut(intensity.ctypes.data, shadow.ctypes.data, rmin.ctypes.data, steps.ctypes.data, 0)
    ab = np.array([alpha, beta], np.float64)
    f = (flow or Flow()).to_c()
    v, o = view.to_c(), _opts(opts)
    _check(t._ctx, K.lib().kt_eht_trace_rays(t._ctx, C.byref(v), C.byref(f), C.c_void_p(ab.ctypes.data), 1,
                                             C.byref(o), C.byref(s), _stream(t, None)), "kt_eht_trace_rays")
    return {"intensity": float(intensity[0]), "captured": bool(shadow[0]), "min_radius": float(rmin[0]),
            "steps": int(steps[0])}


def flux(image: Image, theta_guas: float, tracer: Tracer | None = None, stream=None) -> float:
    """Total flux (Jy) of a device image whose M subtends theta_guas microarcseconds."""
    t = _tracer(tracer)
    inten = image.intensity
    if not hasattr(inten, "data_ptr"):
        inten = t.torch.as_tensor(np.ascontiguousarray(inten, np.float64)).to(f"cuda:{t.device}")
    inten = inten.contiguous()
    jy = C.c_double()
    _check(t._ctx, K.lib().kt_eht_flux(t._ctx, C.c_void_p(inten.data_ptr()), image.size, image.half_width,
                                       theta_guas, C.byref(jy), _stream(t, stream)), "kt_eht_flux")
    return jy.value


def calibrate(view: EhtView, flow: Flow | None, target_jy: float, theta_guas: float, opts: TraceOpts | None = None,
              tracer: Tracer | None = None, stream=None) -> float:
    """The density n0 for which a 48 x 48 render of the view has target_jy janskys.  When
    opts.cancel is set meanwhile, the estimate so far (as eht::calibrate returns it)."""
    t = _tracer(tracer)
    f = (flow or Flow()).to_c()
    v, o = view.to_c(), _opts(opts)
    n0 = C.c_double()
    rc = K.lib().kt_eht_calibrate(t._ctx, C.byref(v), C.byref(f), target_jy, theta_guas, C.byref(o), C.byref(n0),
                                  _stream(t, stream))
    if rc != K.KT_ECANCELLED:
        _check(t._ctx, rc, "kt_eht_calibrate")
    return n0.value


def target_view(target: Target, a: float = 0.9, half_width: float = 16.0, size: int = 128) -> EhtView:
    return EhtView(a=a, inclination=math.radians(target.inclination_deg), half_width=half_width, size=size,
                   mass_solar=target.mass_solar)
