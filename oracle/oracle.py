"""ctypes front end of the FP64 CPU oracle (oracle/kerr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker.  The product (black_simulation_amd)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libkerr_ref.so")

ESCAPED, CAPTURED, DISK, UNFINISHED = 0, 1, 2, 3
BARDEEN, PINHOLE, OBSERVER = 0, 1, 3


class View(C.Structure):
    """Mirror of ko_view (kerr_oracle.h); same fields as kt_view in include/kerr_trace.h."""

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


class Opts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("max_steps", C.c_int)]


class Photon(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("a", "a2", "E", "L", "Q", "aE", "aL", "K", "C1", "A2E2", "twoE",
                                          "twoA2E2")] + [("ks", C.c_int)]


class Result(C.Structure):
    _fields_ = [
        ("cls", C.c_int),
        ("flags", C.c_int),
        ("steps", C.c_int),
        ("theta_inf", C.c_double),
        ("phi_inf", C.c_double),
        ("r_hit", C.c_double),
        ("g", C.c_double),
        ("y_end", C.c_double * 5),
    ]


class Report(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("turns", "r_min", "delay", "b", "q")]


class Consts(C.Structure):
    _fields_ = [("E", C.c_double), ("L", C.c_double), ("Q", C.c_double), ("mu", C.c_double)]


class State(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t", "r", "theta", "phi", "vr", "vtheta")]


class Derivs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "dr", "dtheta", "dphi", "dvr", "dvtheta")]


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


class Shade(C.Structure):
    _fields_ = [("disk_temperature", C.c_double), ("disk_brightness", C.c_double), ("sky", C.c_int),
                ("sky_brightness", C.c_double), ("grid_line", C.c_double), ("sky_image", C.c_void_p),
                ("sky_width", C.c_int), ("sky_height", C.c_int), ("table", C.c_void_p), ("table_n", C.c_int)]


class Flow(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("n0", "T0", "beta", "keplerian", "height", "sin_theta_b")]


class EhtView(C.Structure):
    _fields_ = [("a", C.c_double), ("inclination", C.c_double), ("half_width", C.c_double), ("size", C.c_int),
                ("frequency", C.c_double), ("mass_solar", C.c_double)]


class EhtRay(C.Structure):
    _fields_ = [("intensity", C.c_double), ("min_radius", C.c_double), ("captured", C.c_int), ("steps", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i, P = C.c_double, C.c_int, C.POINTER
        for name, res, args in [
            ("ko_horizon_radius", d, [d]),
            ("ko_inner_horizon_radius", d, [d]),
            ("ko_isco_radius", d, [d, i]),
            ("ko_photon_orbit_radius", d, [d, i]),
            ("ko_marginally_bound_radius", d, [d, i]),
            ("ko_circular_omega", d, [d, d, i]),
            ("ko_derivatives", Derivs, [d, Consts, State]),
            ("ko_radial_potential", d, [d, Consts, d]),
            ("ko_polar_potential", d, [d, Consts, d]),
            ("ko_enforce_constraints", None, [d, Consts, P(State)]),
            ("ko_photon_from_frame", None, [d, d, d, d, P(d), P(d), P(Consts), P(State)]),
            ("ko_zamo_four_velocity", None, [d, d, d, P(d)]),
            ("ko_observer_frame", None, [d, d, d, P(d), P(d)]),
            ("ko_metric_tensor", None, [d, d, d, P(d)]),
            ("ko_camera_basis", None, [d, d, d, P(d)]),
            ("ko_photon_init", None, [P(Photon), d, d, d, d]),
            ("ko_rhs", None, [P(Photon), P(d), P(d)]),
            ("ko_R", d, [P(Photon), d]),
            ("ko_M", d, [P(Photon), d]),
            ("ko_launch_pixel", None, [P(View), i, i, P(Photon), P(d)]),
            ("ko_launch_bardeen", None, [P(View), d, d, P(Photon), P(d)]),
            ("ko_launch_look", None, [P(View), P(d), P(Photon), P(d)]),
            ("ko_observer_u", None, [P(View), P(d)]),
            ("ko_particle_advance", i, [d, d, d, d, P(d), d, d, d, P(Opts), P(i), P(d)]),
            ("ko_particle_from_zamo", None, [d, d, d, P(d), P(d)]),
            ("ko_ks_shift", d, [d, d]),
            ("ko_trace_photon", None, [P(View), P(Opts), P(Photon), P(d), P(Result)]),
            ("ko_trace_photon_report", None, [P(View), P(Opts), P(Photon), P(d), P(Result), P(Report)]),
            ("ko_trace_pixel", None, [P(View), P(Opts), i, i, P(Result)]),
            ("ko_render", C.c_long, [P(View), P(Opts), i, P(C.c_int), i, P(C.c_int8), P(d), P(d), P(d), P(d),
                                     P(C.c_int32)]),
            ("ko_critical_curve", i, [d, d, i, P(d), P(d), i]),
            ("ko_critical_radius", d, [P(d), P(d), i, d]),
            ("ko_pow_m15", d, [d]),
            ("ko_planck_xyz", None, [d, P(d)]),
            ("ko_blackbody_table", None, [i, P(C.c_float)]),
            ("ko_shade_ray", None, [P(Shade), i, d, d, d, d, d, d, P(d)]),
            ("ko_bessel_k", d, [i, d]),
            ("ko_synchrotron_emissivity", d, [d, d, d, d, d]),
            ("ko_planck_nu", d, [d, d]),
            ("ko_plasma_at", None, [P(Flow), d, d, d, P(d)]),
            ("ko_flow_velocity", None, [P(Flow), d, d, d, P(d)]),
            ("ko_eht_trace", None, [P(EhtView), P(Flow), P(Opts), d, d, P(EhtRay)]),
            ("ko_eht_render", None, [P(EhtView), P(Flow), P(Opts), P(d), P(C.c_uint8)]),
            ("ko_eht_flux", d, [P(d), i, d, d]),
            ("ko_eht_calibrate", d, [P(EhtView), P(Flow), P(Opts), d, d]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_view(**kw) -> View:
    v = View()
    defaults = dict(camera_kind=BARDEEN, width=256, height=256, spin=0.9, r_cam=30.0,
                    theta_cam=np.deg2rad(80.0), phi_cam=0.0, half_width=12.0, half_height=12.0,
                    yaw=0.0, pitch=0.0, roll=0.0, fov_deg=60.0, disk_inner=0.0, disk_outer=20.0,
                    disk_prograde=1, disk_enabled=1, escape_radius=1000.0, capture_factor=1.01)
    defaults.update(kw)
    for k, val in defaults.items():
        if k == "u_obs":
            val = (C.c_double * 4)(*[float(x) for x in val])
        elif k == "drift":
            val = (C.c_double * 3)(*[float(x) for x in val])
        elif k == "frame":
            val = (C.c_double * 16)(*[float(x) for x in np.ravel(val)])
        setattr(v, k, val)
    return v


def make_opts(rtol=1e-10, atol=1e-12, max_steps=200000) -> Opts:
    return Opts(rtol, atol, max_steps)


def trace_pixel(view: View, opts: Opts, col: int, row: int) -> Result:
    res = Result()
    lib().ko_trace_pixel(C.byref(view), C.byref(opts), col, row, C.byref(res))
    return res


def render(view: View, opts: Opts, threads: int = 0, stride: int = 1) -> dict:
    """Whole image on `threads` CPU threads (0 = all cores); stride > 1 = sparse sample."""
    if threads <= 0:
        threads = os.cpu_count() or 1
    n = view.width * view.height
    m = (n + stride - 1) // stride if stride > 1 else n
    out = {
        "cls": np.zeros(m, np.int8),
        "theta": np.zeros(m),
        "phi": np.zeros(m),
        "r_hit": np.zeros(m),
        "g": np.zeros(m),
        "steps": np.zeros(m, np.int32),
    }
    got = lib().ko_render(C.byref(view), C.byref(opts), threads, None, stride,
                          out["cls"].ctypes.data_as(C.POINTER(C.c_int8)), _dp(out["theta"]),
                          _dp(out["phi"]), _dp(out["r_hit"]), _dp(out["g"]),
                          out["steps"].ctypes.data_as(C.POINTER(C.c_int32)))
    assert got == m, (got, m)
    return out


def critical_curve(a: float, inclination: float, samples: int):
    cap = 8 * max(samples, 16) * 2 + 4 + samples
    al = np.zeros(cap)
    be = np.zeros(cap)
    n = lib().ko_critical_curve(a, inclination, samples, _dp(al), _dp(be), cap)
    return al[:n].copy(), be[:n].copy()


def critical_radius(alpha: np.ndarray, beta: np.ndarray, psi: float) -> float:
    al = np.ascontiguousarray(alpha, dtype=np.float64)
    be = np.ascontiguousarray(beta, dtype=np.float64)
    return lib().ko_critical_radius(_dp(al), _dp(be), len(al), psi)


REPORT_FIELDS = ("turns", "r_min", "delay", "b", "q")


def trace_params(view: View, opts: Opts, params, report: bool = False) -> dict:
    """Explicit rays: (alpha, beta) rows for a Bardeen view, unit looks for a pinhole view;
    with `report`, also the single-ray report (REPORT_FIELDS, ko_trace_photon_report)."""
    params = np.asarray(params, dtype=np.float64)
    per = 2 if view.camera_kind == BARDEEN else 3
    params = params.reshape(-1, per)
    n = len(params)
    out = {"cls": np.zeros(n, np.int8), "theta": np.full(n, np.nan), "phi": np.full(n, np.nan),
           "r_hit": np.full(n, np.nan), "g": np.full(n, np.nan), "steps": np.zeros(n, np.int32),
           "y_end": np.zeros((n, 5)), "consts": np.zeros((n, 3))}
    if report:
        out.update({k: np.zeros(n) for k in REPORT_FIELDS})
    L = lib()
    for i, row in enumerate(params):
        p = Photon()
        y0 = (C.c_double * 5)()
        if per == 2:
            L.ko_launch_bardeen(C.byref(view), row[0], row[1], C.byref(p), y0)
        else:
            lk = (C.c_double * 3)(*row)
            L.ko_launch_look(C.byref(view), lk, C.byref(p), y0)
        res = Result()
        if report:
            rep = Report()
            L.ko_trace_photon_report(C.byref(view), C.byref(opts), C.byref(p), y0, C.byref(res), C.byref(rep))
            for k in REPORT_FIELDS:
                out[k][i] = getattr(rep, k)
        else:
            L.ko_trace_photon(C.byref(view), C.byref(opts), C.byref(p), y0, C.byref(res))
        out["cls"][i] = res.cls
        out["steps"][i] = res.steps
        out["y_end"][i] = list(res.y_end)
        out["consts"][i] = (p.E, p.L, p.Q)
        if res.cls == ESCAPED:
            out["theta"][i] = res.theta_inf
            out["phi"][i] = res.phi_inf
        if res.cls == DISK:
            out["r_hit"][i] = res.r_hit
            out["g"][i] = res.g
    return out


def planck_xyz(T: float) -> np.ndarray:
    out = np.zeros(3)
    lib().ko_planck_xyz(T, _dp(out))
    return out


def blackbody_table(n: int = 1024) -> np.ndarray:
    t = np.zeros(3 * n, np.float32)
    lib().ko_blackbody_table(n, t.ctypes.data_as(C.POINTER(C.c_float)))
    return t.reshape(n, 3)


def shade(rays: dict, r_in, r_out, disk_temperature=10000.0, disk_brightness=1.0, sky=1, sky_brightness=1.0,
          grid_line=0.02, sky_image=None, table_n=1024) -> np.ndarray:
    """Colour of every ray (rays: cls, theta_inf, phi_inf, r_hit, g as numpy arrays; r_in,
    r_out per ray or scalars) -> (n, 3) float64."""
    table = blackbody_table(table_n)
    img = None if sky_image is None else np.ascontiguousarray(sky_image, np.float32)
    s = Shade(disk_temperature, disk_brightness, sky, sky_brightness, grid_line,
              None if img is None else img.ctypes.data, 0 if img is None else img.shape[1],
              0 if img is None else img.shape[0], table.ctypes.data, table_n)
    n = len(rays["cls"])
    r_in = np.broadcast_to(np.asarray(r_in, np.float64), (n,))
    r_out = np.broadcast_to(np.asarray(r_out, np.float64), (n,))
    out = np.zeros((n, 3))
    buf = np.zeros(3)
    L = lib()
    for q in range(n):
        L.ko_shade_ray(C.byref(s), int(rays["cls"][q]) & 7, float(rays["theta_inf"][q]), float(rays["phi_inf"][q]),
                       float(rays["r_hit"][q]), float(rays["g"][q]), float(r_in[q]), float(r_out[q]), _dp(buf))
        out[q] = buf
    return out


def flow(n0=1e5, T0=4e11, beta=10.0, keplerian=0.9, height=1.0, sin_theta_b=0.8165) -> Flow:
    return Flow(n0, T0, beta, keplerian, height, sin_theta_b)


def eht_view(a=0.9, inclination=0.3, half_width=16.0, size=128, frequency=230e9, mass_solar=6.5e9) -> EhtView:
    return EhtView(a, inclination, half_width, size, frequency, mass_solar)


def plasma_at(f: Flow, r: float, theta: float, a: float) -> np.ndarray:
    out = np.zeros(3)
    lib().ko_plasma_at(C.byref(f), r, theta, a, _dp(out))
    return out


def flow_velocity(f: Flow, a: float, r: float, theta: float) -> np.ndarray:
    out = np.zeros(4)
    lib().ko_flow_velocity(C.byref(f), a, r, theta, _dp(out))
    return out


def eht_trace(v: EhtView, f: Flow, opts: Opts, alpha: float, beta: float) -> EhtRay:
    r = EhtRay()
    lib().ko_eht_trace(C.byref(v), C.byref(f), C.byref(opts), alpha, beta, C.byref(r))
    return r


def eht_render(v: EhtView, f: Flow, opts: Opts):
    n = v.size * v.size
    inten = np.zeros(n)
    shadow = np.zeros(n, np.uint8)
    lib().ko_eht_render(C.byref(v), C.byref(f), C.byref(opts), _dp(inten), shadow.ctypes.data_as(C.POINTER(C.c_uint8)))
    return inten.reshape(v.size, v.size), shadow.reshape(v.size, v.size)


def eht_flux(intensity: np.ndarray, half_width: float, theta_guas: float) -> float:
    a = np.ascontiguousarray(intensity, np.float64)
    return lib().ko_eht_flux(_dp(a), a.shape[0], half_width, theta_guas)


def eht_calibrate(v: EhtView, f: Flow, opts: Opts, target_jy: float, theta_guas: float) -> float:
    return lib().ko_eht_calibrate(C.byref(v), C.byref(f), C.byref(opts), target_jy, theta_guas)


PARTICLE_MOVING, PARTICLE_PLUNGED, PARTICLE_ESCAPED, PARTICLE_UNFINISHED = 0, 1, 2, 3


def particle_from_zamo(a: float, r: float, theta: float, v) -> np.ndarray:
    """kerr::launchFromZamo, mu = 1: (E, L, Q, vr, vmu)."""
    out = np.zeros(5)
    vv = np.ascontiguousarray(v, np.float64)
    lib().ko_particle_from_zamo(a, r, theta, _dp(vv), _dp(out))
    return out


def particle_advance(a, E, L, Q, state, dt, r_stop, r_escape, opts: Opts, h=None):
    """(fate, new state (r, mu, phi, vr, vmu), DP5 attempts); with h given (the step to start
    with, 0 = the default) also the controller's next step: (fate, state, steps, h)."""
    st = np.ascontiguousarray(state, np.float64).copy()
    steps = C.c_int()
    hh = None if h is None else C.c_double(h)
    fate = lib().ko_particle_advance(a, E, L, Q, _dp(st), dt, r_stop, r_escape, C.byref(opts), C.byref(steps),
                                     None if hh is None else C.byref(hh))
    if h is None:
        return fate, st, steps.value
    return fate, st, steps.value, hh.value
