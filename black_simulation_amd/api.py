"""Python face of the MI355X Kerr tracer, mirroring the reference's own entry points.

    render(view)            ~ eht::render(View, Flow, threads, cancel)   (src/eht.h:140)
    trace_ray(view, a, b)   ~ eht::traceRay(View, Flow, alpha, beta)     (src/eht.h:130)
    trace_looks(view, looks)~ trajectory::traceRay(a, r, th, ph, frame, look, scene)
                                                                        (src/trajectory.h:118)
    horizon_radius, isco_radius, photon_orbit_radius, circular_omega      (src/kerr.h:39-45, 71)
    critical_curve, critical_radius                                       (src/eht.h:66-68)

All work goes through libkerr_trace.so (include/kerr_trace.h); torch is used only for
device buffers and streams.  There is no CPU fallback: without the library and a GPU
these functions raise.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Iterable, Sequence

import numpy as np

from . import _lib as K

OUTPUT_FIELDS = ("cls", "theta_inf", "phi_inf", "r_hit", "g", "steps")
FLOAT_FIELDS = ("theta_inf", "phi_inf", "r_hit", "g")  # float64, or float32 (kt_rays_out *_f32)
# the single-ray report (trajectory::RayReport, src/trajectory.h:97-113), float64; any of
# them in `out` selects the FP64 report trace
REPORT_FIELDS = ("turns", "r_min", "delay", "b", "q")


@dataclasses.dataclass
class View:
    """One camera view (kt_view).  Bardeen: the image plane of eht::View; pinhole: the
    simulator's Camera (src/camera.h:20-29) in the ZAMO frame; observer: that camera
    carried by the observer of 4-velocity u_obs (Kerr-Schild, e.g. from
    geodesic_four_velocity; inside the horizon too), or by the ZAMO boosted by `drift`
    (Camera::frame, src/camera.cpp:23-40), or given its frame outright: `frame`, 16
    values e_a^mu (rows: 4-velocity, r, theta, phi axes; Kerr-Schild), as uCamPhoton
    carries Camera::frame (src/renderer.cpp:345-357); observer_frame builds one."""

    camera_kind: int = K.CAMERA_BARDEEN
    width: int = 256
    height: int = 256
    spin: float = 0.9
    r_cam: float = 30.0
    theta_cam: float = math.radians(80.0)
    phi_cam: float = 0.0
    half_width: float = 12.0
    half_height: float = 12.0
    yaw: float = 0.0
    pitch: float = 0.0
    roll: float = 0.0
    fov_deg: float = 60.0
    disk_inner: float = 0.0
    disk_outer: float = 20.0
    disk_prograde: int = 1
    disk_enabled: int = 1
    escape_radius: float = 1000.0
    capture_factor: float = 1.01
    u_obs: tuple = (0.0, 0.0, 0.0, 0.0)
    drift: tuple = (0.0, 0.0, 0.0)
    frame: tuple = (0.0,) * 16

    def to_c(self) -> K.KtView:
        v = K.KtView()
        for f in dataclasses.fields(self):
            val = getattr(self, f.name)
            if f.name == "u_obs":
                val = (C.c_double * 4)(*[float(x) for x in val])
            elif f.name == "drift":
                val = (C.c_double * 3)(*[float(x) for x in val])
            elif f.name == "frame":
                val = (C.c_double * 16)(*[float(x) for x in np.ravel(val)])
            setattr(v, f.name, val)
        return v

    @property
    def rays(self) -> int:
        return self.width * self.height


@dataclasses.dataclass
class TraceOpts:
    precision: int = K.PREC_FP64
    mode: int = K.MODE_AUTO
    use_surrogate: int = 0
    max_steps: int = 200000
    rtol: float = 1e-10
    atol: float = 1e-12
    rtol32: float = 1e-6
    atol32: float = 1e-7
    restep_drift: float = 1e-6
    block: int = 0
    waves_per_cu: int = 0
    steps_per_dequeue: int = 0
    gate_prob: float = 0.999
    gate_ring: float = 0.01
    waves_per_simd: int = 0
    # EHT render / calibrate: a ctypes.c_int another thread may set to cancel (eht::render's
    # cancel flag); None: not cancellable
    cancel: object = None
    # 1: the closed-form solution answers the rays it covers (KT_FLAG_CLOSED_FORM), the
    # integrator the rest; with use_surrogate, it replaces the values the network answered.
    # 2: the same with FP32 elliptic integrals (FP64 for the rays FP32 leaves unsure)
    closed_form: int = 0

    def to_c(self) -> K.KtTraceOpts:
        o = K.KtTraceOpts()
        for f in dataclasses.fields(self):
            if f.name == "cancel":
                o.cancel = C.addressof(self.cancel) if self.cancel is not None else None
            else:
                setattr(o, f.name, getattr(self, f.name))
        return o


def _check(ctx, rc: int, what: str):
    if rc != K.KT_OK:
        msg = K.lib().kt_last_error_string(ctx)
        raise RuntimeError(f"{what} failed (status {rc}): {msg.decode() if msg else ''}")


class Tracer:
    """A library context (kt_ctx) on one GPU."""

    def __init__(self, device: int = 0):
        import torch  # device memory and streams only

        if not torch.cuda.is_available():
            raise RuntimeError("no GPU: the Kerr tracer runs only on the HIP device")
        self.torch = torch
        self.device = device
        self._ctx = C.c_void_p()
        opts = K.KtDeviceOpts(device)
        rc = K.lib().kt_create(C.byref(opts), C.byref(self._ctx))
        if rc != K.KT_OK:
            raise RuntimeError(f"kt_create failed with status {rc}")

    def close(self):
        if self._ctx:
            K.lib().kt_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- device outputs -------------------------------------------------------------
    def alloc_outputs(self, n: int, f32: bool = False, report: bool = False) -> dict:
        """Device outputs for n rays; f32: the sky angles, disk radius and g in float (each
        the FP64 result rounded once), 21 instead of 37 bytes per ray; report: also the
        single-ray report fields (REPORT_FIELDS, float64)."""
        t = self.torch
        dev = t.device("cuda", self.device)
        real = t.float32 if f32 else t.float64
        out = {
            "cls": t.empty(n, dtype=t.int8, device=dev),
            "theta_inf": t.empty(n, dtype=real, device=dev),
            "phi_inf": t.empty(n, dtype=real, device=dev),
            "r_hit": t.empty(n, dtype=real, device=dev),
            "g": t.empty(n, dtype=real, device=dev),
            "steps": t.empty(n, dtype=t.int32, device=dev),
        }
        if report:
            out.update({k: t.empty(n, dtype=t.float64, device=dev) for k in REPORT_FIELDS})
        return out

    @staticmethod
    def _set_out(s: K.KtRaysOut, name: str, ptr, is_f32: bool):
        """one output field of kt_rays_out: the float pointer for FP32 buffers"""
        setattr(s, name + "_f32" if is_f32 and name in FLOAT_FIELDS else name, ptr)

    def _out_struct(self, out: dict) -> K.KtRaysOut:
        s = K.KtRaysOut()
        for name in OUTPUT_FIELDS:
            tens = out.get(name)
            if tens is not None:
                self._set_out(s, name, tens.data_ptr(), tens.dtype == self.torch.float32)
        for name in REPORT_FIELDS:
            tens = out.get(name)
            if tens is not None:
                if tens.dtype != self.torch.float64 or not tens.is_contiguous():
                    raise ValueError(f"report output {name}: contiguous float64 needed")
                setattr(s, name, tens.data_ptr())
        s.on_device = 1
        return s

    def trace_views(self, views: Sequence[View], opts: TraceOpts | None = None, out: dict | None = None,
                    stream=None, report: bool = False) -> dict:
        """Trace every pixel of `views` in one launch; returns device tensors (views
        concatenated).  report: fresh outputs include the single-ray report (FP64 only)."""
        n = sum(v.rays for v in views)
        if out is None:
            out = self.alloc_outputs(n, report=report)
        arr = (K.KtView * len(views))(*[v.to_c() for v in views])
        o = (opts or TraceOpts()).to_c()
        s = self._out_struct(out)
        st = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        rc = K.lib().kt_trace_views(self._ctx, arr, len(views), C.byref(o), C.byref(s), C.c_void_p(st.cuda_stream))
        _check(self._ctx, rc, "kt_trace_views")
        return out

    def host_outputs(self, n: int, pinned: bool = True, f32: bool = False) -> dict:
        """Host buffers for n rays' outputs to reuse across trace_views_host calls: CPU
        tensors, page-locked when pinned (the copy back then runs at the link's DMA rate
        instead of through a pageable staging copy); f32 as alloc_outputs."""
        t = self.torch
        real = t.float32 if f32 else t.float64
        dt = {"cls": t.int8, "theta_inf": real, "phi_inf": real, "r_hit": real, "g": real, "steps": t.int32}
        return {k: t.empty(n, dtype=dt[k], pin_memory=pinned) for k in OUTPUT_FIELDS}

    def trace_views_host(self, views: Sequence[View], opts: TraceOpts | None = None, out: dict | None = None,
                         report: bool = False) -> dict:
        """As trace_views, into host buffers: the library copies the outputs back over PCIe
        and synchronises before returning.  `out`: reusable buffers (host_outputs, or
        contiguous numpy arrays of the output dtypes, with REPORT_FIELDS too for the
        report); None: fresh numpy arrays (with the report when `report`)."""
        n = sum(v.rays for v in views)
        dt = {"cls": np.int8, "theta_inf": np.float64, "phi_inf": np.float64, "r_hit": np.float64,
              "g": np.float64, "steps": np.int32, **{k: np.float64 for k in REPORT_FIELDS}}
        if out is None:
            out = {k: np.empty(n, dt[k]) for k in OUTPUT_FIELDS + (REPORT_FIELDS if report else ())}
        s = K.KtRaysOut()
        for name in OUTPUT_FIELDS + REPORT_FIELDS:
            if name not in out:
                if name in REPORT_FIELDS:
                    continue
                raise ValueError(f"host output {name} missing")
            buf = out[name]
            if isinstance(buf, np.ndarray):
                is_f32 = name in FLOAT_FIELDS and buf.dtype == np.float32
                if (buf.dtype != dt[name] and not is_f32) or buf.size < n or not buf.flags["C_CONTIGUOUS"]:
                    raise ValueError(f"host output {name}: {n} contiguous {np.dtype(dt[name])} values needed")
                self._set_out(s, name, buf.ctypes.data, is_f32)
            else:
                is_f32 = name in FLOAT_FIELDS and buf.dtype == self.torch.float32
                if buf.device.type != "cpu" or buf.numel() < n or not buf.is_contiguous() or \
                        (buf.element_size() != np.dtype(dt[name]).itemsize and not is_f32):
                    raise ValueError(f"host output {name}: {n} contiguous CPU values of {np.dtype(dt[name])} needed")
                self._set_out(s, name, buf.data_ptr(), is_f32)
        s.on_device = 0
        arr = (K.KtView * len(views))(*[v.to_c() for v in views])
        o = (opts or TraceOpts()).to_c()
        st = self.torch.cuda.current_stream(self.device)
        _check(self._ctx, K.lib().kt_trace_views(self._ctx, arr, len(views), C.byref(o), C.byref(s),
                                                 C.c_void_p(st.cuda_stream)), "kt_trace_views")
        return out

    def trace_params(self, view: View, params, opts: TraceOpts | None = None, out: dict | None = None,
                     stream=None, report: bool = False) -> dict:
        """Explicit rays: (alpha, beta) rows for a Bardeen view, unit looks for a pinhole view.
        report: fresh outputs include the single-ray report (FP64 only)."""
        t = self.torch
        per = 2 if view.camera_kind == K.CAMERA_BARDEEN else 3
        p = t.as_tensor(np.asarray(params, dtype=np.float64).reshape(-1, per)) if not isinstance(params, t.Tensor) \
            else params.to(t.float64).reshape(-1, per)
        p = p.to(f"cuda:{self.device}").contiguous()
        n = p.shape[0]
        if out is None:
            out = self.alloc_outputs(n, report=report)
        o = (opts or TraceOpts()).to_c()
        s = self._out_struct(out)
        v = view.to_c()
        st = stream if stream is not None else t.cuda.current_stream(self.device)
        rc = K.lib().kt_trace_rays(self._ctx, C.byref(v), C.c_void_p(p.data_ptr()), n, C.byref(o), C.byref(s),
                                   C.c_void_p(st.cuda_stream))
        _check(self._ctx, rc, "kt_trace_rays")
        self._keep = p  # alive until the stream has used it
        return out

    def autotune(self, views: Sequence[View], opts: TraceOpts | None = None) -> TraceOpts:
        """kt_autotune: time launch shapes on `views`, return the fastest (also cached in
        the context for later calls whose shape fields are 0)."""
        o = (opts or TraceOpts()).to_c()
        arr = (K.KtView * len(views))(*[v.to_c() for v in views])
        _check(self._ctx, K.lib().kt_autotune(self._ctx, arr, len(views), C.byref(o)), "kt_autotune")
        res = dataclasses.replace(opts or TraceOpts())
        for f in ("mode", "block", "waves_per_cu", "steps_per_dequeue", "waves_per_simd"):
            setattr(res, f, getattr(o, f))
        return res

    def plan(self, views: Sequence[View], opts: TraceOpts | None = None, out: dict | None = None) -> "Plan":
        """A batch captured once into a hipGraph (kt_plan_create), replayed by Plan.launch()."""
        n = sum(v.rays for v in views)
        if out is None:
            out = self.alloc_outputs(n)
        arr = (K.KtView * len(views))(*[v.to_c() for v in views])
        o = (opts or TraceOpts()).to_c()
        s = self._out_struct(out)
        handle = C.c_void_p()
        _check(self._ctx, K.lib().kt_plan_create(self._ctx, arr, len(views), C.byref(o), C.byref(s),
                                                 C.byref(handle)), "kt_plan_create")
        return Plan(self, handle, out)

    # -- the learned surrogate ------------------------------------------------------------
    def surrogate_init(self, seed: int = 1):
        _check(self._ctx, K.lib().kt_surrogate_init(self._ctx, seed), "kt_surrogate_init")

    def surrogate_train(self, **kw) -> dict:
        """Generate data with the integrator, train on the GPU, judge the held-out set."""
        o = K.KtSurrogateOpts()
        K.lib().kt_surrogate_opts_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        r = K.KtSurrogateReport()
        _check(self._ctx, K.lib().kt_surrogate_train(self._ctx, C.byref(o), C.byref(r)), "kt_surrogate_train")
        return {f: getattr(r, f) for f, _ in K.KtSurrogateReport._fields_}

    def surrogate_set_net(self, net: int):
        """Chooses the network (NET_WIDE or NET_FAST); its parameters are dropped."""
        _check(self._ctx, K.lib().kt_surrogate_set_net(self._ctx, net), "kt_surrogate_set_net")

    def surrogate_net(self) -> int:
        return K.lib().kt_surrogate_get_net(self._ctx)

    def surrogate_save(self) -> np.ndarray:
        n = K.lib().kt_surrogate_net_param_count(self.surrogate_net())
        p = np.zeros(n, np.float32)
        _check(self._ctx, K.lib().kt_surrogate_save(self._ctx, p.ctypes.data, n), "kt_surrogate_save")
        return p

    def surrogate_load(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(self._ctx, K.lib().kt_surrogate_load(self._ctx, p.ctypes.data, p.size), "kt_surrogate_load")

    def surrogate_predict(self, inputs):
        """Network outputs for normalised inputs (n x 5) -> (n x 8), on the device."""
        t = self.torch
        x = t.as_tensor(np.asarray(inputs, np.float32)).to(f"cuda:{self.device}").contiguous()
        y = t.empty((x.shape[0], 8), dtype=t.float32, device=x.device)
        st = t.cuda.current_stream(self.device)
        _check(self._ctx, K.lib().kt_surrogate_predict(self._ctx, x.data_ptr(), x.shape[0], y.data_ptr(),
                                                       C.c_void_p(st.cuda_stream)), "kt_surrogate_predict")
        return y

    def surrogate_gradients(self, inputs, cls, targets):
        t = self.torch
        dev = f"cuda:{self.device}"
        x = t.as_tensor(np.asarray(inputs, np.float32)).to(dev).contiguous()
        c = t.as_tensor(np.asarray(cls, np.int8)).to(dev).contiguous()
        y = t.as_tensor(np.asarray(targets, np.float32)).to(dev).contiguous()
        g = t.empty(K.lib().kt_surrogate_net_param_count(self.surrogate_net()), dtype=t.float32, device=dev)
        loss = C.c_double()
        t.cuda.synchronize()
        _check(self._ctx, K.lib().kt_surrogate_gradients(self._ctx, x.data_ptr(), c.data_ptr(), y.data_ptr(),
                                                         x.shape[0], g.data_ptr(), C.byref(loss)),
               "kt_surrogate_gradients")
        return g, loss.value

    def surrogate_adam_step(self, grads, lr: float, t: int):
        self.torch.cuda.synchronize()
        _check(self._ctx, K.lib().kt_surrogate_adam_step(self._ctx, grads.data_ptr(), lr, t), "kt_surrogate_adam_step")

    def surrogate_domain(self) -> dict:
        """What the loaded network covers (kt_surrogate_get_domain)."""
        d = K.KtSurrogateDomain()
        _check(self._ctx, K.lib().kt_surrogate_get_domain(self._ctx, C.byref(d)), "kt_surrogate_get_domain")
        return {f: getattr(d, f) for f, _ in K.KtSurrogateDomain._fields_}

    def surrogate_set_domain(self, **kw):
        """Changes fields of the network's domain (kt_surrogate_set_domain)."""
        d = K.KtSurrogateDomain()
        _check(self._ctx, K.lib().kt_surrogate_get_domain(self._ctx, C.byref(d)), "kt_surrogate_get_domain")
        for k, v in kw.items():
            if k not in dict(K.KtSurrogateDomain._fields_):
                raise KeyError(k)
            setattr(d, k, float(v))
        _check(self._ctx, K.lib().kt_surrogate_set_domain(self._ctx, C.byref(d)), "kt_surrogate_set_domain")

    def surrogate_shadow_radius(self, a: float, inclination: float) -> float:
        return K.lib().kt_surrogate_shadow_radius(self._ctx, a, inclination)

    def surrogate_table_radius(self, a: float, inclination: float, psi: float) -> float:
        return K.lib().kt_surrogate_table_radius(self._ctx, a, inclination, psi)

    # -- HDR colour ----------------------------------------------------------------------
    def shade(self, views: Sequence[View], rays: dict, sky: int = K.SKY_GRID, disk_temperature: float = 1e4,
              disk_brightness: float = 1.0, sky_brightness: float = 1.0, grid_line: float = 0.02, sky_image=None,
              stream=None):
        """kt_shade_views: RGBA (half floats, n x 4, on the device) of the traced `rays`
        (device outputs of trace_views on the same views).  sky_image: (H, W, 3) float32,
        equirectangular, row 0 at latitude -90 deg."""
        t = self.torch
        n = sum(v.rays for v in views)
        rgba = t.empty((n, 4), dtype=t.float16, device=f"cuda:{self.device}")
        o = K.KtShadeOpts()
        K.lib().kt_shade_opts_default(C.byref(o))
        o.sky, o.disk_temperature, o.disk_brightness = sky, disk_temperature, disk_brightness
        o.sky_brightness, o.grid_line = sky_brightness, grid_line
        img = None
        if sky_image is not None:
            img = t.as_tensor(np.ascontiguousarray(sky_image, np.float32)).to(rgba.device)
            o.sky_image, o.sky_width, o.sky_height = img.data_ptr(), img.shape[1], img.shape[0]
        arr = (K.KtView * len(views))(*[v.to_c() for v in views])
        s = self._out_struct(rays)
        st = stream if stream is not None else t.cuda.current_stream(self.device)
        _check(self._ctx, K.lib().kt_shade_views(self._ctx, arr, len(views), C.byref(s), C.byref(o),
                                                 C.c_void_p(rgba.data_ptr()), C.c_void_p(st.cuda_stream)),
               "kt_shade_views")
        self._keep_img = img
        return rgba

    def stats(self) -> dict:
        s = K.KtStats()
        _check(self._ctx, K.lib().kt_last_stats(self._ctx, C.byref(s)), "kt_last_stats")
        return {f: getattr(s, f) for f, _ in K.KtStats._fields_}


class Plan:
    """A captured batch (kt_plan); launch() replays its hipGraph on a stream."""

    def __init__(self, tracer: Tracer, handle, out: dict):
        self.tracer, self._h, self.out = tracer, handle, out

    def launch(self, stream=None) -> dict:
        st = stream if stream is not None else self.tracer.torch.cuda.current_stream(self.tracer.device)
        _check(self.tracer._ctx, K.lib().kt_plan_launch(self._h, C.c_void_p(st.cuda_stream)), "kt_plan_launch")
        return self.out

    def stats(self) -> dict:
        """kt_plan_stats: counters and time of the last replay (waits for it)."""
        s = K.KtStats()
        _check(self.tracer._ctx, K.lib().kt_plan_stats(self._h, C.byref(s)), "kt_plan_stats")
        return {f: getattr(s, f) for f, _ in K.KtStats._fields_}

    def close(self):
        if self._h:
            K.lib().kt_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_numpy(out: dict) -> dict:
    return {k: v.cpu().numpy() for k, v in out.items()}


_default_tracer: Tracer | None = None


def _tracer() -> Tracer:
    global _default_tracer
    if _default_tracer is None:
        _default_tracer = Tracer()
    return _default_tracer


def render(view: View, opts: TraceOpts | None = None) -> dict:
    """eht::render counterpart: every pixel of one view, as host numpy arrays reshaped to
    (height, width), row-major from the top."""
    out = _tracer().trace_views_host([view], opts)
    return {k: v.reshape(view.height, view.width) for k, v in out.items()}


def trace_ray(view: View, alpha: float, beta: float, opts: TraceOpts | None = None, report: bool = False) -> dict:
    """eht::traceRay counterpart for one (alpha, beta) of a Bardeen view; with `report`
    also the RayReport fields turns, r_min, delay, b, q (FP64)."""
    out = to_numpy(_tracer().trace_params(view, [[alpha, beta]], opts, report=report))
    return {k: v[0].item() for k, v in out.items()}


def trace_looks(view: View, looks: Iterable[Sequence[float]], opts: TraceOpts | None = None,
                report: bool = False) -> dict:
    """trajectory::traceRay counterpart: unit look vectors of a pinhole view's ZAMO frame;
    with `report` also its RayReport fields (src/trajectory.h:97-113): turns, r_min, delay,
    b, q (FP64)."""
    return to_numpy(_tracer().trace_params(view, np.asarray(list(looks), dtype=np.float64), opts, report=report))


def blackbody_table(n: int = 1024) -> np.ndarray:
    """kt_blackbody_table: (n, 3) linear-sRGB blackbody colours over log10 T in [2.5, 5.3]."""
    t = np.zeros(3 * n, np.float32)
    rc = K.lib().kt_blackbody_table(n, t.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != K.KT_OK:
        raise RuntimeError(f"kt_blackbody_table failed with status {rc}")
    return t.reshape(n, 3)


def observer_frame(a: float, r: float, theta: float, u) -> np.ndarray:
    """kerr::observerFrame (src/kerr.cpp:432-456) of the observer of 4-velocity u
    (Kerr-Schild) at (r, theta): 4 x 4, rows e_0 = u normalised, then the r, theta, phi axes."""
    uu = np.ascontiguousarray(u, np.float64)
    e = np.zeros(16)
    rc = K.lib().kt_observer_frame(a, r, theta, uu.ctypes.data_as(C.POINTER(C.c_double)),
                                   e.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != K.KT_OK:
        raise ValueError("observer_frame: u must be a future-directed timelike vector")
    return e.reshape(4, 4)


def geodesic_four_velocity(a: float, E: float, L: float, Q: float, r: float, theta: float, radial_sign: int = -1,
                           polar_sign: int = 1) -> tuple:
    """kt_geodesic_four_velocity: (u^t, u^r, u^theta, u^phi) in Kerr-Schild components of a
    free-falling massive observer (kerr::fourVelocity, src/kerr.cpp:414-418)."""
    u = (C.c_double * 4)()
    rc = K.lib().kt_geodesic_four_velocity(a, E, L, Q, r, theta, radial_sign, polar_sign, u)
    if rc != K.KT_OK:
        raise ValueError(f"no geodesic with E={E}, L={L}, Q={Q} at r={r}, theta={theta}")
    return tuple(u)


def horizon_radius(a: float) -> float:
    return K.lib().kt_horizon_radius(a)


def isco_radius(a: float, prograde: bool = True) -> float:
    return K.lib().kt_isco_radius(a, int(prograde))


def photon_orbit_radius(a: float, prograde: bool = True) -> float:
    return K.lib().kt_photon_orbit_radius(a, int(prograde))


def circular_omega(a: float, r: float, prograde: bool = True) -> float:
    return K.lib().kt_circular_omega(a, r, int(prograde))


def critical_curve(a: float, inclination: float, samples: int = 400):
    cap = 16 * max(samples, 16) + 8 + samples
    al = np.zeros(cap)
    be = np.zeros(cap)
    dp = C.POINTER(C.c_double)
    n = K.lib().kt_critical_curve(a, inclination, samples, al.ctypes.data_as(dp), be.ctypes.data_as(dp), cap)
    return al[:n].copy(), be[:n].copy()


def critical_radius(alpha, beta, psi: float) -> float:
    al = np.ascontiguousarray(alpha, dtype=np.float64)
    be = np.ascontiguousarray(beta, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    return K.lib().kt_critical_radius(al.ctypes.data_as(dp), be.ctypes.data_as(dp), len(al), psi)


def shadow_radius(a: float, inclination: float, samples: int = 400, directions: int = 72) -> float:
    """Mean critical radius over `directions` angles about the image origin
    (as src/eht_view.cpp:101 sizes its field)."""
    al, be = critical_curve(a, inclination, samples)
    return float(np.mean([critical_radius(al, be, 2 * math.pi * k / directions) for k in range(directions)]))
