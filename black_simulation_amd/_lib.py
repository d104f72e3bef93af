"""ctypes binding of libkerr_trace.so (the C ABI declared in include/kerr_trace.h).

The shared library is the product; this module only loads it.  It raises if the
library is missing -- there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libkerr_trace.so")

KT_OK, KT_EINVAL, KT_EHIP, KT_ECANCELLED, KT_ENOMEM, KT_EUNSUPPORTED, KT_ENODEVICE = range(7)
ESCAPED, CAPTURED, DISK, UNFINISHED = 0, 1, 2, 3
FLAG_CLOSED_FORM, FLAG_FP64_RESTEP, FLAG_SURROGATE = 0x08, 0x10, 0x20
CAMERA_BARDEEN, CAMERA_PINHOLE, CAMERA_OBSERVER = 0, 1, 3
NET_WIDE, NET_FAST = 0, 1
PARTICLE_MOVING, PARTICLE_PLUNGED, PARTICLE_ESCAPED, PARTICLE_UNFINISHED = 0, 1, 2, 3
PREC_FP64, PREC_MIXED, PREC_FP32 = 0, 1, 2
MODE_PERSISTENT, MODE_ONESHOT, MODE_AUTO = 0, 1, 2


class KtView(C.Structure):
    _fields_ = [
        ("camera_kind", C.c_int),
        ("width", C.c_int),
        ("height", C.c_int),
        ("spin", C.c_double),
        ("r_cam", C.c_double),
        ("theta_cam", C.c_double),
        ("phi_cam", C.c_double),
        ("half_width", C.c_double),
        ("half_height", C.c_double),
        ("yaw", C.c_double),
        ("pitch", C.c_double),
        ("roll", C.c_double),
        ("fov_deg", C.c_double),
        ("disk_inner", C.c_double),
        ("disk_outer", C.c_double),
        ("disk_prograde", C.c_int),
        ("disk_enabled", C.c_int),
        ("escape_radius", C.c_double),
        ("capture_factor", C.c_double),
        ("u_obs", C.c_double * 4),
        ("drift", C.c_double * 3),
        ("frame", C.c_double * 16),
    ]


class KtTraceOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("mode", C.c_int),
        ("use_surrogate", C.c_int),
        ("max_steps", C.c_int),
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("rtol32", C.c_double),
        ("atol32", C.c_double),
        ("restep_drift", C.c_double),
        ("block", C.c_int),
        ("waves_per_cu", C.c_int),
        ("steps_per_dequeue", C.c_int),
        ("gate_prob", C.c_double),
        ("gate_ring", C.c_double),
        ("waves_per_simd", C.c_int),
        ("cancel", C.c_void_p),
        ("closed_form", C.c_int),
    ]


class KtRaysOut(C.Structure):
    _fields_ = [
        ("cls", C.c_void_p),
        ("theta_inf", C.c_void_p),
        ("phi_inf", C.c_void_p),
        ("r_hit", C.c_void_p),
        ("g", C.c_void_p),
        ("steps", C.c_void_p),
        ("on_device", C.c_int),
        ("theta_inf_f32", C.c_void_p),
        ("phi_inf_f32", C.c_void_p),
        ("r_hit_f32", C.c_void_p),
        ("g_f32", C.c_void_p),
        ("turns", C.c_void_p),
        ("r_min", C.c_void_p),
        ("delay", C.c_void_p),
        ("b", C.c_void_p),
        ("q", C.c_void_p),
    ]


class KtDeviceOpts(C.Structure):
    _fields_ = [("device", C.c_int)]


class KtStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("integrator_ms", C.c_double),
        ("surrogate_ms", C.c_double),
        ("rays", C.c_longlong),
        ("steps", C.c_longlong),
        ("fallback_rays", C.c_longlong),
        ("restep_rays", C.c_longlong),
        ("fp64_steps", C.c_longlong),
        ("block", C.c_int),
        ("waves_per_cu", C.c_int),
        ("steps_per_dequeue", C.c_int),
        ("grid", C.c_int),
        ("waves_per_simd", C.c_int),
        ("mode", C.c_int),
    ]


class KtSurrogateOpts(C.Structure):
    _fields_ = [
        ("n_train", C.c_longlong),
        ("n_holdout", C.c_longlong),
        ("epochs", C.c_int),
        ("batch", C.c_int),
        ("lr", C.c_double),
        ("w_sky", C.c_double),
        ("w_disk", C.c_double),
        ("band", C.c_double),
        ("band_factor", C.c_double),
        ("seed", C.c_ulonglong),
        ("label_precision", C.c_int),
        ("disk_outer", C.c_double),
        ("escape_radius", C.c_double),
        ("capture_factor", C.c_double),
        ("verbose", C.c_int),
        ("net", C.c_int),
        ("edge_band", C.c_double),
        ("edge_copies", C.c_int),
    ]


class KtSurrogateDomain(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("spin_min", "spin_max", "inclination_min", "inclination_max",
                                          "r_cam_min", "r_cam_max", "extent", "disk_outer", "escape_radius",
                                          "capture_factor")]


class KtSurrogateReport(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("datagen_seconds", "train_seconds", "eval_seconds", "first_loss",
                                          "final_loss", "train_gemm_tflops")] + [("steps", C.c_longlong)] + \
               [(n, C.c_double) for n in ("class_agreement", "gated_fraction", "gated_class_agreement",
                                          "defl_median", "defl_p99", "defl_p999", "g_rel_median", "g_rel_p99",
                                          "rhit_rel_median", "rhit_rel_p99")]


class KtEhtFlow(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("n0", "T0", "beta", "keplerian", "height", "sin_theta_b")]


class KtEhtView(C.Structure):
    _fields_ = [("a", C.c_double), ("inclination", C.c_double), ("half_width", C.c_double), ("size", C.c_int),
                ("frequency", C.c_double), ("mass_solar", C.c_double)]


class KtEhtOut(C.Structure):
    _fields_ = [("intensity", C.c_void_p), ("shadow", C.c_void_p), ("min_radius", C.c_void_p), ("steps", C.c_void_p),
                ("on_device", C.c_int)]


class KtParticles(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("r", "mu", "phi", "vr", "vmu", "E", "L", "Q", "fate", "steps", "h")] + \
        [("on_device", C.c_int)]


class KtShadeOpts(C.Structure):
    _fields_ = [
        ("disk_temperature", C.c_double),
        ("disk_brightness", C.c_double),
        ("sky", C.c_int),
        ("sky_brightness", C.c_double),
        ("grid_line", C.c_double),
        ("sky_image", C.c_void_p),
        ("sky_width", C.c_int),
        ("sky_height", C.c_int),
    ]


SKY_BLACK, SKY_GRID, SKY_IMAGE = 0, 1, 2

# every entry point include/kerr_trace.h declares: name -> (restype, argtypes)
_P = C.POINTER
SIGNATURES = {
    "kt_trace_opts_default": (None, [_P(KtTraceOpts)]),
    "kt_view_default": (None, [_P(KtView)]),
    "kt_create": (C.c_int, [_P(KtDeviceOpts), _P(C.c_void_p)]),
    "kt_destroy": (C.c_int, [C.c_void_p]),
    "kt_last_error_string": (C.c_char_p, [C.c_void_p]),
    "kt_last_stats": (C.c_int, [C.c_void_p, _P(KtStats)]),
    "kt_trace_views": (C.c_int, [C.c_void_p, _P(KtView), C.c_int, _P(KtTraceOpts), _P(KtRaysOut), C.c_void_p]),
    "kt_trace_views_noalloc": (C.c_int, [C.c_void_p, _P(KtView), C.c_int, _P(KtTraceOpts), _P(KtRaysOut),
                                         C.c_void_p]),
    "kt_trace_rays": (C.c_int, [C.c_void_p, _P(KtView), C.c_void_p, C.c_longlong, _P(KtTraceOpts), _P(KtRaysOut),
                                C.c_void_p]),
    "kt_autotune": (C.c_int, [C.c_void_p, _P(KtView), C.c_int, _P(KtTraceOpts)]),
    "kt_plan_create": (C.c_int, [C.c_void_p, _P(KtView), C.c_int, _P(KtTraceOpts), _P(KtRaysOut), _P(C.c_void_p)]),
    "kt_plan_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kt_plan_stats": (C.c_int, [C.c_void_p, _P(KtStats)]),
    "kt_plan_destroy": (C.c_int, [C.c_void_p]),
    "kt_surrogate_opts_default": (None, [_P(KtSurrogateOpts)]),
    "kt_surrogate_param_count": (C.c_longlong, []),
    "kt_surrogate_net_param_count": (C.c_longlong, [C.c_int]),
    "kt_surrogate_set_net": (C.c_int, [C.c_void_p, C.c_int]),
    "kt_surrogate_get_net": (C.c_int, [C.c_void_p]),
    "kt_surrogate_init": (C.c_int, [C.c_void_p, C.c_ulonglong]),
    "kt_surrogate_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong]),
    "kt_surrogate_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong]),
    "kt_surrogate_train": (C.c_int, [C.c_void_p, _P(KtSurrogateOpts), _P(KtSurrogateReport)]),
    "kt_surrogate_predict": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p]),
    "kt_surrogate_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                         _P(C.c_double)]),
    "kt_surrogate_adam_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int]),
    "kt_surrogate_table_radius": (C.c_double, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "kt_surrogate_shadow_radius": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    "kt_surrogate_domain_default": (None, [_P(KtSurrogateDomain)]),
    "kt_surrogate_get_domain": (C.c_int, [C.c_void_p, _P(KtSurrogateDomain)]),
    "kt_surrogate_set_domain": (C.c_int, [C.c_void_p, _P(KtSurrogateDomain)]),
    "kt_shade_opts_default": (None, [_P(KtShadeOpts)]),
    "kt_shade_views": (C.c_int, [C.c_void_p, _P(KtView), C.c_int, _P(KtRaysOut), _P(KtShadeOpts), C.c_void_p,
                                 C.c_void_p]),
    "kt_blackbody_table": (C.c_int, [C.c_int, _P(C.c_float)]),
    "kt_eht_flow_default": (None, [_P(KtEhtFlow)]),
    "kt_eht_view_default": (None, [_P(KtEhtView)]),
    "kt_eht_render": (C.c_int, [C.c_void_p, _P(KtEhtView), _P(KtEhtFlow), _P(KtTraceOpts), _P(KtEhtOut), C.c_void_p]),
    "kt_eht_trace_rays": (C.c_int, [C.c_void_p, _P(KtEhtView), _P(KtEhtFlow), C.c_void_p, C.c_longlong,
                                    _P(KtTraceOpts), _P(KtEhtOut), C.c_void_p]),
    "kt_eht_flux": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, _P(C.c_double), C.c_void_p]),
    "kt_eht_calibrate": (C.c_int, [C.c_void_p, _P(KtEhtView), _P(KtEhtFlow), C.c_double, C.c_double,
                                   _P(KtTraceOpts), _P(C.c_double), C.c_void_p]),
    "kt_eht_angular_scale": (C.c_double, [C.c_double, C.c_double]),
    "kt_observer_frame": (C.c_int, [C.c_double, C.c_double, C.c_double, _P(C.c_double), _P(C.c_double)]),
    "kt_geodesic_four_velocity": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_int, C.c_int, _P(C.c_double)]),
    "kt_particles_advance": (C.c_int, [C.c_void_p, C.c_double, _P(KtParticles), C.c_longlong, C.c_double,
                                       C.c_double, C.c_double, _P(KtTraceOpts), C.c_void_p]),
    "kt_particle_from_zamo": (C.c_int, [C.c_double, C.c_double, C.c_double, _P(C.c_double), _P(C.c_double)]),
    "kt_horizon_radius": (C.c_double, [C.c_double]),
    "kt_isco_radius": (C.c_double, [C.c_double, C.c_int]),
    "kt_photon_orbit_radius": (C.c_double, [C.c_double, C.c_int]),
    "kt_circular_omega": (C.c_double, [C.c_double, C.c_double, C.c_int]),
    "kt_critical_curve": (C.c_int, [C.c_double, C.c_double, C.c_int, _P(C.c_double), _P(C.c_double), C.c_int]),
    "kt_critical_radius": (C.c_double, [_P(C.c_double), _P(C.c_double), C.c_int, C.c_double]),
}

_lib = None


def lib():
    """Load libkerr_trace.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               "g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
