"""Headline benchmark: Mrays/s rendering 1920x1080 frames of Kerr a=0.998, surrogate +
integrator fallback against the plain FP64 integrator on the same frames.

Workload (default): the 120-frame camera orbit at r = 40 M (phi_k = 2 pi k / 120,
inclination 20 -> 85 deg), Bardeen plane +-1.5 mean shadow radii, disk r_ISCO..20 M, escape 1000 M,
capture 1.01 r+.  One step = one full frame (2,073,600 rays), replayed from a hipGraph
plan.  --path surrogate: the MLP (by default the fast network, 5-128x4-8 with hard SiLU)
answers rays with max class probability > 0.999 lying outside +-3% of the critical
curve, the integrator the rest (FP64 DP5 at rtol 1e-6; --fallback mixed: the FP32/FP64
one), all in one graph; --path plain: the FP64 integrator (rtol 1e-10); --path mixed:
the FP32/FP64 integrator (rtol 1e-6).  Both the chosen path and the plain FP64 integrator are timed on the same frames, and
the chosen path's outputs are judged against the FP64 render of every frame.
--workload c2 times the single a=0.998 frame at r = 50 M, i = 85 deg (disk to 40 M)
instead; the surrogate does not cover that scene.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--path surrogate|plain|mixed]
                    [--workload orbit|c2] [--height 1080|2160] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s 1080p a=0.998: surrogate+fallback vs plain integrator @≤1e-4 rad"

# Floating-point operations of one DP5 attempt of the integrator (kt_device.hpp), each
# +, -, *, /, sqrt counted once and an FMA as two:
#   6 right-hand sides x 25 (one division each)                  = 150
#   stage sums (per component 3 + 5 + 7 + 9 + 11 + 11, x 5)       = 230
#   error estimate (5 x 12) and max-norm (5 x 4 + 1 division)     =  81
#   controller (x^-1/5 by three Newton steps, safety, h update)   =  30
#   constraint projection (R, M, two zone tests, two sqrt)         =  31
OPS_PER_ATTEMPT = 522
# One DP5 attempt of a massive particle in coordinate time (kt_particles.hip), counted the
# same way: 6 right-hand sides x 40 (three divisions each: the inflow factor, L / sin^2 and
# 1 / (dt/dlambda)) = 240, stage sums 230, error estimate and max-norm 81, controller 30,
# projection 35, and the right-hand side re-evaluated after each accepted step 40
OPS_PER_PARTICLE_ATTEMPT = 656

# HBM bytes per launch of each workload's dominant kernel, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# (the factors a 1 GiB device copy reads) from rocprofv3 --pmc runs of the same command;
# the summaries (commands included) are the named files under profiles/
PROFILED_TRAFFIC = {
    ("orbit/closed", 1080, "k_closed_form"): (0.377e6 + 76.73e6, "profiles/k_closed_form_orbit_1080p.md"),
    ("orbit/surrogate", 1080, "k_mlp_fast"): (1.220e6 + 75.30e6, "profiles/k_mlp_fast_orbit_1080p.md"),
    ("orbit/surrogate", 1080, "k_trace_oneshot"): (1.203e6 + 3.885e6,
                                                   "profiles/k_trace_oneshot_fallback_orbit_1080p.md"),
    ("infall", 1080, "k_trace_persistent"): (0.374e6 + 149.1e6, "profiles/k_trace_persistent_plainks_infall_1080p.md"),
    ("eht", 1024, "k_eht"): (2.468e6 + 12.13e6, "profiles/k_eht_m87_1024.md"),
    ("debris", 1 << 22, "k_advance"): (339.2e6 + 899.9e6, "profiles/k_advance_debris_4m.md"),
}


def profiled_traffic(workload, size, kernel):
    """(bytes, source) of the profiled launch, or (None, None) for other workloads/sizes"""
    for (w, n, k), v in PROFILED_TRAFFIC.items():
        if w == workload and n == size and kernel.startswith(k):
            return v
    return None, None


# Operations per ray of the closed-form solution (kt_closed.hpp) over the rays of the 1080p
# orbit frames, counted by its Python restatement (tests/closed_form.py `count_ops`: + - * /
# sqrt once, FMA twice, each iteration of the special functions as run): the FP64 form, and
# the FP32 pass with the FP64 one for the rays it leaves unsure
CLOSED_FORM_OPS_PER_RAY = {"fp64": 293, "fp32": 276}


def mlp_flop_per_ray(hidden: int) -> int:
    """multiply-adds x 2 of the network 5 -> H -> H -> H -> H -> 8: 399,872 at H = 256 (the
    wide network), 101,632 at H = 128 (the fast one)"""
    return 2 * (5 * hidden + 3 * hidden * hidden + hidden * 8)


FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0


def cpu_baseline(view, seconds_target: float = 15.0, more=()) -> dict:
    """The FP64 oracle (same DP5, rtol 1e-10) on this host's cores over a sample of the
    workload sized for ~15 s: every `stride`-th pixel of one frame, or, when the whole frame
    takes less than that, whole frames (`view`, then the views in `more`) until it is spent."""
    from oracle import oracle as O

    cores = os.cpu_count() or 1

    def oracle_view(v):
        return O.make_view(**{k: getattr(v, k) for k in (
            "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
            "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
            "escape_radius", "capture_factor", "u_obs", "drift", "frame")})

    ov = oracle_view(view)
    opts = O.make_opts()
    stride = 4099
    t0 = time.perf_counter()
    out = O.render(ov, opts, threads=cores, stride=stride)
    dt = time.perf_counter() - t0
    rate = len(out["cls"]) / max(dt, 1e-6)
    want = max(1, int(rate * seconds_target))
    stride = max(1, (view.width * view.height) // want)
    if stride == 1 and more:
        n, dt, used = 0, 0.0, 0
        for v in [view, *more]:
            t0 = time.perf_counter()
            n += len(O.render(oracle_view(v), opts, threads=cores)["cls"])
            dt += time.perf_counter() - t0
            used += 1
            if dt >= seconds_target:
                break
        return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
                "sample": f"{used} whole {view.width}x{view.height} frames of the workload ({n} rays, "
                          f"{dt:.1f} s, FP64 DP5 rtol 1e-10, {cores} threads)"}
    t0 = time.perf_counter()
    out = O.render(ov, opts, threads=cores, stride=stride)
    dt = time.perf_counter() - t0
    n = len(out["cls"])
    return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th pixel of one {view.width}x{view.height} frame of the workload ({n} rays, "
                      f"{dt:.1f} s, FP64 DP5 rtol 1e-10, {cores} threads)"}


def c1_timing(tr, torch) -> dict:
    """The CPU-runnable configuration in full (a = 0.9, r = 30 M, i = 80 deg, 256^2 Bardeen
    pixels, FP64 DP5 rtol 1e-10): the GPU's FP64 path and its closed form (HIP events around
    the whole call, best of 5) and the oracle on every pixel over the host's cores; the
    classes agree (the parity tests check every output)."""
    import numpy as np

    import black_simulation_amd as bsa
    from black_simulation_amd.configs import c1_view
    from oracle import oracle as O

    v = c1_view()
    stream = torch.cuda.current_stream()

    def best_of_5(opts):
        out = tr.trace_views([v], opts)
        best = 1e30
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            tr.trace_views([v], opts, out=out)
            e1.record(stream)
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best, out["cls"].cpu().numpy()

    best, got = best_of_5(bsa.TraceOpts(precision=bsa.PREC_FP64))
    best_cf, got_cf = best_of_5(bsa.TraceOpts(closed_form=2))
    cores = os.cpu_count() or 1
    ov = O.make_view(**{k: getattr(v, k) for k in (
        "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
        "disk_inner", "disk_outer", "disk_prograde", "disk_enabled", "escape_radius", "capture_factor")})
    t0 = time.perf_counter()
    want = O.render(ov, O.make_opts(), threads=cores)
    cpu_s = time.perf_counter() - t0
    return {"rays": v.rays, "gpu_fp64_ms": round(best, 3), "gpu_fp64_mrays_s": round(v.rays / best / 1e3, 2),
            "gpu_closed_form_ms": round(best_cf, 4), "gpu_closed_form_mrays_s": round(v.rays / best_cf / 1e3, 1),
            "cpu_port_s": round(cpu_s, 3), "cpu_port_mrays_s": round(v.rays / cpu_s / 1e6, 4), "cpu_threads": cores,
            "classes_equal": bool(np.array_equal(got & 7, want["cls"])),
            "closed_form_classes_equal": bool(np.array_equal(got_cf & 7, want["cls"]))}


def _ref_lib():
    """oracle/_ref/libkerr_ref.so: the reference's own sources compiled as they are (with
    oracle/ref_shim.cpp), or None where it was not built"""
    import ctypes as C

    from oracle import oracle as O

    return C.CDLL(O.REF_LIB_PATH) if os.path.exists(O.REF_LIB_PATH) else None


def _timed_pool(one, n, seconds_target):
    """one(q) over every stride-th q < n from a thread pool of the host's cores (ctypes
    releases the GIL), the stride sized for ~seconds_target; (rays, seconds, stride, cores)"""
    from concurrent.futures import ThreadPoolExecutor

    cores = os.cpu_count() or 1

    def run(stride):
        idx = list(range(stride // 2, n, stride))
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(one, idx, chunksize=16))
        return len(idx), time.perf_counter() - t0

    k, dt = run(max(1, n // (64 * cores)))
    stride = max(1, int(n / max(1.0, k / dt * seconds_target)))
    k, dt = run(stride)
    return k, dt, stride, cores


def infall_cpu_reference(view, seconds_target: float = 10.0):
    """The reference's own trajectory::traceRay (src/trajectory.cpp:214-396, its RK4 and
    step rule) from the falling observer's frame (kerr::observerFrame), on the pixels' look
    directions (camera.cpp:6-21 basis, trace.frag:1494-1496 pixel mapping) of the same
    frame; None when oracle/_ref was not built."""
    import ctypes as C

    import numpy as np

    lib = _ref_lib()
    if lib is None:
        return None
    dp = C.POINTER(C.c_double)
    lib.ref_trajectory_trace_u.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, C.c_int,
                                           C.c_double, C.c_double, C.c_double, dp]
    lib.ref_trajectory_trace_u.restype = None
    cy, sy, cp, sp, cr, sr = (math.cos(view.yaw), math.sin(view.yaw), math.cos(view.pitch), math.sin(view.pitch),
                              math.cos(view.roll), math.sin(view.roll))
    f0, u0, r0 = np.array([-1.0, 0.0, 0.0]), np.array([0.0, -1.0, 0.0]), np.array([0.0, 0.0, 1.0])
    yawed, right = cy * f0 + sy * r0, cy * r0 - sy * f0
    fwd = cp * yawed + sp * u0
    up = np.cross(right, fwd)
    B = (cr * right + sr * up, cr * up - sr * right, fwd)
    tanh_fov = math.tan(0.5 * view.fov_deg * math.pi / 180.0)
    W, H = view.width, view.height
    u = (C.c_double * 4)(*view.u_obs)
    r_in = view.disk_inner if view.disk_inner > 0 else __import__("black_simulation_amd").isco_radius(view.spin, True)

    def one(q):
        row, col = divmod(q, W)
        x = ((col + 0.5) - 0.5 * W) / (0.5 * H) * tanh_fov
        y = (((H - row) - 0.5) - 0.5 * H) / (0.5 * H) * tanh_fov
        d = x * B[0] + y * B[1] + B[2]
        d /= np.linalg.norm(d)
        look = (C.c_double * 3)(*d)
        out = (C.c_double * 11)()
        lib.ref_trajectory_trace_u(view.spin, view.r_cam, view.theta_cam, view.phi_cam, u, look,
                                   1 if view.disk_enabled else 0, r_in, view.disk_outer, view.escape_radius, out)
        return out[0]

    k, dt, stride, cores = _timed_pool(one, W * H, seconds_target)
    return {"value": round(k / dt / 1e6, 5), "unit": "Mrays/s", "cores": cores, "kind": "reference",
            "sample": f"every {stride}th pixel of one {W}x{H} frame ({k} rays, {dt:.1f} s): the reference's "
                      f"trajectory::traceRay from the observer's frame (its RK4), {cores} threads"}


def bardeen_cpu_reference(view, seconds_target: float = 10.0):
    """The reference's own trajectory::traceRay (src/trajectory.cpp:214-396: RK4 with its
    step rule at tol 0.01, linear disk crossing) from the camera's ZAMO frame, on every
    stride-th pixel of one frame of a Bardeen-plane workload, from a thread pool of the
    host's cores; None when oracle/_ref was not built.  Pixel (alpha, beta) becomes the
    look direction whose photon has L = -alpha sin(i) and Q = beta^2 + (alpha^2 - a^2)
    cos^2(i) (ZAMO-frame components of p: sqrt(R)/sqrt(Sigma Delta), sign(beta)
    sqrt(Theta)/sqrt(Sigma), L/sqrt(g_phiphi), reversed and normalised).  Its classes, disk
    radii and g are set beside the FP64 oracle's on the same pixels; its g, relative to the
    ZAMO's unit photon energy, times the photon's conserved energy E is g seen from infinity."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from oracle import oracle as O

    lib = _ref_lib()
    if lib is None:
        return None
    dp = C.POINTER(C.c_double)
    lib.ref_trajectory_trace.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp, C.c_int,
                                         C.c_double, C.c_double, C.c_double, dp]
    lib.ref_trajectory_trace.restype = None
    a, r, th, W, H = view.spin, view.r_cam, view.theta_cam, view.width, view.height
    st, ct = math.sin(th), math.cos(th)
    S, D = r * r + a * a * ct * ct, r * r - 2 * r + a * a
    gpp = ((r * r + a * a) ** 2 - a * a * D * st * st) * st * st / S
    r_in = view.disk_inner if view.disk_inner > 0 else __import__("black_simulation_amd").isco_radius(view.spin, True)
    n = W * H

    def one(q):
        row, col = divmod(q, W)
        al = -view.half_width + (col + 0.5) * 2.0 * view.half_width / W
        be = view.half_height - (row + 0.5) * 2.0 * view.half_height / H
        L, Q = -al * st, be * be + (al * al - a * a) * ct * ct
        R = (r * r + a * a - a * L) ** 2 - D * (Q + (L - a) ** 2)
        Th = Q - ct * ct * (L * L / (st * st) - a * a)
        p = np.array([math.sqrt(max(R, 0.0)) / math.sqrt(S * D), math.copysign(math.sqrt(max(Th, 0.0)), be) / math.sqrt(S),
                      L / math.sqrt(gpp)])
        look = (C.c_double * 3)(*(-p / np.linalg.norm(p)))
        out = (C.c_double * 11)()
        lib.ref_trajectory_trace(a, r, th, view.phi_cam, look, 1 if view.disk_enabled else 0, r_in, view.disk_outer,
                                 view.escape_radius, out)
        return out[0], out[1], out[2] * out[8]  # g seen by the ZAMO times E: g seen from infinity

    cores = os.cpu_count() or 1

    def run(stride):
        idx = list(range(0, n, stride))  # the oracle's sparse sample: idx % stride == 0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            res = list(ex.map(one, idx, chunksize=16))
        return res, time.perf_counter() - t0

    res, dt = run(max(1, n // (64 * cores)))
    stride = max(1, int(n / max(1.0, len(res) / dt * seconds_target)))
    res, dt = run(stride)
    k = len(res)
    end = np.array([e for e, _, _ in res], dtype=np.int64)
    cls = np.select([end == 0, end == 1, end == 2], [O.ESCAPED, O.CAPTURED, O.DISK], O.UNFINISHED)
    ov = O.make_view(**{f: getattr(view, f) for f in (
        "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
        "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
        "escape_radius", "capture_factor")})
    want = O.render(ov, O.make_opts(), threads=cores, stride=stride)
    both = (cls == O.DISK) & (want["cls"] == O.DISK)
    r_rel = np.abs(np.array([x for _, x, _ in res])[both] / want["r_hit"][both] - 1.0)
    g_rel = np.abs(np.array([x for _, _, x in res])[both] / want["g"][both] - 1.0)
    pct = lambda x: [float("%.3g" % v) for v in np.percentile(x, [50, 99])] if x.size else None  # noqa: E731
    return {"value": round(k / dt / 1e6, 5), "unit": "Mrays/s", "cores": cores, "kind": "reference",
            "sample": f"every {stride}th pixel of one {W}x{H} frame ({k} rays, {dt:.1f} s): the reference's "
                      f"trajectory::traceRay from the camera's ZAMO frame (its RK4), {cores} threads",
            "vs_fp64_oracle": {"class_agreement": round(float((cls == want["cls"]).mean()), 5),
                               "unfinished": int((cls == O.UNFINISHED).sum()),
                               "r_hit_rel_p50_p99": pct(r_rel), "g_rel_p50_p99": pct(g_rel)}}


def percentile(xs, q):
    s = sorted(xs)
    return s[min(len(s) - 1, int(q * (len(s) - 1) + 0.5))]


def timed_frames(orbit, warm, frames, torch):
    """Replays each frame once, frames alternating over the orbit's streams; returns
    (wall seconds, per-frame graph ms from each plan's HIP events)."""
    for k in warm:
        orbit.render(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in frames:
        orbit.render(k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [orbit.stats(k)["kernel_ms"] for k in frames]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", choices=["closed", "surrogate", "plain", "mixed"], default="closed")
    ap.add_argument("--closed-form", type=int, choices=[0, 1], default=1,
                    help="--path surrogate: 1 = the closed form verifies the network's escaped and disk answers "
                         "and gives their values; 0 = the network's own values")
    ap.add_argument("--closed-precision", choices=["fp32", "fp64"], default="fp32",
                    help="--path closed: the elliptic integrals in FP32 (FP64 for the rays FP32 leaves unsure; "
                         "the default) or in FP64 throughout")
    ap.add_argument("--no-alt", action="store_true",
                    help="--path closed: do not time the learned surrogate's paths beside it")
    ap.add_argument("--workload", choices=["orbit", "c2", "c3", "eht", "infall", "debris"], default="orbit")
    ap.add_argument("--particles", type=int, default=1 << 22, help="--workload debris: particles")
    ap.add_argument("--size", type=int, default=1024, help="--workload eht: pixels across")
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--weights", default=None,
                    help="surrogate parameters (PATH.f32 and PATH.json); default the fast network, "
                         "black_simulation_amd/data/surrogate_fast (the wide one: .../data/surrogate)")
    ap.add_argument("--streams", type=int, default=3)
    ap.add_argument("--lean", action="store_true",
                    help="for profilers: no autotuning, no plain-FP64 timing, no accuracy pass")
    ap.add_argument("--host-outputs", action="store_true",
                    help="also time the path with host output buffers (PCIe copy included, one stream)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="profiler calibration only: five device copies of this many MiB, then exit")
    ap.add_argument("--mode", choices=["auto", "oneshot", "persistent"], default="auto",
                    help="integrator kernel (auto: the autotuner's pick, or persistent with --lean)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fallback", choices=["mixed", "fp64"], default="fp64",
                    help="the integrator of --path surrogate: FP64 (default) at rtol 1e-6, or the mixed FP32/FP64 "
                         "one; --rtol sets its tolerance")
    ap.add_argument("--gate-ring", type=float, default=0.0,
                    help="surrogate gate: least |rho / rho_c - 1| of an answered ray (0: the library's default)")
    ap.add_argument("--gate-prob", type=float, default=0.0,
                    help="surrogate gate: least class probability of an answered ray (0: the library's default)")
    ap.add_argument("--rtol", type=float, default=0.0,
                    help="DP5 tolerance of the chosen path (FP64 rtol for --path plain, the FP32 bulk's for mixed "
                         "and surrogate; atol = rtol / 10); the reference render stays FP64 at rtol 1e-10")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the tracer runs on a single MI355X (rays are independent; how a "
              f"launch would shard across GPUs is described, not built); --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.path == "surrogate" and args.workload == "c2":
        print("the surrogate is trained for the orbit scene (disk to 20 M); use --workload orbit", file=sys.stderr)
        sys.exit(2)
    if args.workload == "infall" and args.path != "plain":
        print("observer cameras trace in FP64: --workload infall runs --path plain", file=sys.stderr)
        args.path = "plain"

    import torch

    if args.copy_mib:  # a plain copy of known size, to scale FETCH_SIZE / WRITE_SIZE readings
        n = args.copy_mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}), flush=True)
        return

    if args.workload == "eht":
        eht_bench(args, torch)
        return
    if args.workload == "debris":
        debris_bench(args, torch)
        return

    import black_simulation_amd as bsa
    from black_simulation_amd import orbit as O
    from black_simulation_amd.configs import C5_FRAMES, c2_view, c3_views, c5_view, infall_view
    from black_simulation_amd.train_surrogate import FAST_WEIGHTS, load_into

    width = args.height * 16 // 9
    if args.workload == "orbit":
        views = [c5_view(k, width, args.height) for k in range(C5_FRAMES)]
        frames = []
        for i in range(args.steps):  # evenly spaced, consecutive frames on different streams
            k = (i * C5_FRAMES) // args.steps % C5_FRAMES
            if frames and k % args.streams == frames[-1] % args.streams:
                k = (k + 1) % C5_FRAMES
            frames.append(k)
        warm = [(C5_FRAMES // 2 + i) % C5_FRAMES for i in range(args.warmup)]
    else:  # one frame (a view, or the 64-view batch), a copy per stream so that replays can overlap
        one = ([c2_view(width, args.height)] if args.workload == "c2" else
               [infall_view(width, args.height)] if args.workload == "infall" else c3_views())
        views = [one] * args.streams
        frames = [i % args.streams for i in range(args.steps)]
        warm = [i % args.streams for i in range(args.warmup)]
    frame_views = [v if isinstance(v, list) else [v] for v in views]
    rays = sum(v.rays for v in frame_views[0])

    tr = bsa.Tracer(0)
    tune = (lambda o: o) if args.lean else (lambda o: tr.autotune(frame_views[frames[len(frames) // 2]], o))
    mode = {"auto": bsa.MODE_AUTO, "oneshot": bsa.MODE_ONESHOT, "persistent": bsa.MODE_PERSISTENT}[args.mode]
    plain = tune(bsa.TraceOpts(precision=bsa.PREC_FP64, mode=mode))
    def surrogate_opts(closed_form):
        rtol = args.rtol or 1e-6
        if args.fallback == "fp64":
            o = tune(bsa.TraceOpts(precision=bsa.PREC_FP64, mode=mode, rtol=rtol, atol=rtol / 10))
        else:
            o = tune(bsa.TraceOpts(precision=bsa.PREC_MIXED, mode=mode,
                                   **({"rtol32": args.rtol, "atol32": args.rtol / 10} if args.rtol else {})))
        o.use_surrogate = 1
        o.closed_form = closed_form
        if args.gate_ring:
            o.gate_ring = args.gate_ring
        if args.gate_prob:
            o.gate_prob = args.gate_prob
        return o

    alt = []  # (name, opts, answered flag): the learned surrogate's paths beside --path closed
    if args.path == "closed":
        # the closed form answers every ray it covers; the FP64 integrator (rtol 1e-10, or
        # --rtol) the rays it declines
        opts = tune(bsa.TraceOpts(precision=bsa.PREC_FP64, mode=mode,
                                  **({"rtol": args.rtol, "atol": args.rtol / 10} if args.rtol else {})))
        opts.closed_form = 2 if args.closed_precision == "fp32" else 1
        if args.workload == "orbit" and not args.no_alt and not args.lean:
            other = bsa.TraceOpts(**{f: getattr(opts, f) for f in ("precision", "mode", "rtol", "atol", "block",
                                                                    "waves_per_cu", "steps_per_dequeue",
                                                                    "waves_per_simd")})
            other.closed_form = 1 if opts.closed_form == 2 else 2
            alt = [(f"closed_form_{'fp64' if other.closed_form == 1 else 'fp32'}", other, bsa.FLAG_CLOSED_FORM)]
            load_into(tr, args.weights or FAST_WEIGHTS)
            alt += [("learned_surrogate_closed_form_values", surrogate_opts(opts.closed_form), bsa.FLAG_SURROGATE),
                    ("learned_surrogate_network_values", surrogate_opts(0), bsa.FLAG_SURROGATE)]
    elif args.path == "plain":
        opts = plain
        if args.rtol:  # FP64 at another tolerance, judged against FP64 at rtol 1e-10
            opts = tune(bsa.TraceOpts(precision=bsa.PREC_FP64, mode=mode, rtol=args.rtol, atol=args.rtol / 10))
    elif args.path == "surrogate":
        load_into(tr, args.weights or FAST_WEIGHTS)
        opts = surrogate_opts(args.closed_form)
    else:
        opts = tune(bsa.TraceOpts(precision=bsa.PREC_MIXED, mode=mode,
                                  **({"rtol32": args.rtol, "atol32": args.rtol / 10} if args.rtol else {})))
    tr_net = tr.surrogate_net() if (args.path == "surrogate" or alt) else None
    answered_flag = bsa.FLAG_CLOSED_FORM if args.path == "closed" else bsa.FLAG_SURROGATE

    # the chosen path, timed
    orb = O.Orbit(tr, views, opts, streams=args.streams)
    wall, lat = timed_frames(orb, warm, frames, torch)
    wall_frames = len(frames)
    value = rays * wall_frames / wall / 1e6
    # the plain FP64 integrator on the same frames, timed the same way
    ref = O.Orbit(tr, views, plain, streams=args.streams) if (args.path != "plain" or args.rtol) and not args.lean \
        else orb
    if ref is not orb:
        wall_p, lat_p = timed_frames(ref, warm, frames, torch)
    else:
        wall_p, lat_p = wall, lat
    value_p = rays * wall_frames / wall_p / 1e6
    # the learned surrogate's paths beside the closed form, timed the same way
    alt_runs = []
    for name, aopts, aflag in alt:
        aorb = O.Orbit(tr, views, aopts, streams=args.streams)
        awall, alat = timed_frames(aorb, warm, frames, torch)
        alt_runs.append((name, aopts, aflag, aorb, rays * wall_frames / awall / 1e6, alat))

    # accuracy of each path against the FP64 render of every timed frame
    acc = None
    alt_acc = {}
    if ref is not orb:
        paths = [("main", orb, answered_flag)] + [(r[0], r[3], r[2]) for r in alt_runs]
        tally = {name: {"agree": 0, "answered": 0, "n": 0, "counts": {}, "defl": [], "defl_a": [], "g": [],
                        "g_a": [], "r": [], "r_a": []} for name, _, _ in paths}
        for k in sorted(set(frames)):
            theirs = ref.render(k)
            ref.synchronize()
            for name, porb, pflag in paths:
                mine = porb.render(k)
                porb.synchronize()
                e = O.errors(mine, theirs, pflag)
                t = tally[name]
                t["agree"] += int(e["agree"].sum())
                t["answered"] += int(e["answered"].sum())
                t["n"] += rays
                for c, d in e["counts"].items():
                    tc = t["counts"].setdefault(c, {"rays": 0, "answered": 0, "answered_agree": 0})
                    for f in tc:
                        tc[f] += d[f]
                for key, src in (("defl", "defl"), ("defl_a", "defl_answered"), ("g", "g_rel"),
                                 ("g_a", "g_rel_answered"), ("r", "r_rel"), ("r_a", "r_rel_answered")):
                    t[key].append(e[src])
        q = lambda xs: [float("%.3g" % v) for v in O.quantiles(torch.cat(xs))]  # noqa: E731

        def summary(name, v):
            t = tally[name]
            counts = t["counts"]
            a = {"frames": len(set(frames)), "class_agreement": round(t["agree"] / t["n"], 6),
                 "answered_fraction": round(t["answered"] / t["n"], 4),
                 # the frames' class mix (FP64 classes) and, per class, how many rays the path
                 # answered without the integrator and how many of those got the right class
                 "rays_by_class": counts,
                 "answered_class_agreement": round(sum(c["answered_agree"] for c in counts.values())
                                                   / max(1, t["answered"]), 6),
                 "deflection_rad_p50_p99_p999": q(t["defl"]), "deflection_rad_answered_p50_p99_p999": q(t["defl_a"]),
                 "g_rel_p50_p99_p999": q(t["g"]), "g_rel_answered_p50_p99_p999": q(t["g_a"]),
                 "r_hit_rel_p50_p99_p999": q(t["r"]), "r_hit_rel_answered_p50_p99_p999": q(t["r_a"])}
            # the target judges what the path answered without the integrator: its escaped
            # rays' deflection (none answered: not met), every ray's class, and the speed
            # against plain FP64; the integrator paths are judged on all their rays
            if name != "main" or args.path in ("closed", "surrogate"):
                d50, _, d999 = a["deflection_rad_answered_p50_p99_p999"]
                a["target_met"] = bool(counts["escaped"]["answered"] > 0 and v >= 20 * value_p and d50 <= 1e-4
                                       and d999 <= 1e-3 and a["class_agreement"] >= 0.999)
            else:
                d50, _, d999 = a["deflection_rad_p50_p99_p999"]
                a["target_met"] = bool(v >= 20 * value_p and d50 <= 1e-4 and d999 <= 1e-3
                                       and a["class_agreement"] >= 0.999)
            return a

        acc = summary("main", value)
        for name, aopts, aflag, aorb, avalue, alat in alt_runs:
            alt_acc[name] = {"value": round(avalue, 3), "speedup_vs_plain_fp64": round(avalue / value_p, 3),
                             "frame_ms_p50": round(percentile(alat, 0.5), 3),
                             "gate": {"prob": aopts.gate_prob, "ring": aopts.gate_ring,
                                      "escaped_and_disk_verified_by_closed_form": bool(aopts.closed_form)}
                             if aopts.use_surrogate else None,
                             "accuracy_vs_fp64": summary(name, avalue)}
    for r in alt_runs:
        r[3].close()

    # the same frames through host buffers (the library copies the outputs back over PCIe)
    # through page-locked buffers allocated once (host_outputs), and through fresh pageable
    # numpy arrays per frame
    host_value = host_value_pageable = host_value_f32 = None
    if args.host_outputs:
        k0 = frames[len(frames) // 2]
        hb = tr.host_outputs(rays, pinned=True)
        tr.trace_views_host(frame_views[k0], opts, out=hb)
        t0 = time.perf_counter()
        for k in frames:
            tr.trace_views_host(frame_views[k], opts, out=hb)
        host_value = rays * len(frames) / (time.perf_counter() - t0) / 1e6
        hb = tr.host_outputs(rays, pinned=True, f32=True)  # 21 B/ray: float angles, radius and g
        tr.trace_views_host(frame_views[k0], opts, out=hb)
        t0 = time.perf_counter()
        for k in frames:
            tr.trace_views_host(frame_views[k], opts, out=hb)
        host_value_f32 = rays * len(frames) / (time.perf_counter() - t0) / 1e6
        tr.trace_views_host(frame_views[k0], opts)
        t0 = time.perf_counter()
        for k in frames[: max(1, len(frames) // 4)]:
            tr.trace_views_host(frame_views[k], opts)
        host_value_pageable = rays * max(1, len(frames) // 4) / (time.perf_counter() - t0) / 1e6

    # per-kernel attribution (untimed pass without graphs: events around each kernel group;
    # the last GPU work of the path, so a kernel trace of this run ends with its dispatches)
    integ_ms, sur_ms, attempts, fp64_att, fallback = [], [], [], [], []
    for k in frames[:: max(1, len(frames) // 24)]:
        tr.trace_views(frame_views[k], opts, out=orb.out)
        st = tr.stats()
        integ_ms.append(st["integrator_ms"])
        sur_ms.append(st["surrogate_ms"])
        attempts.append(st["steps"])
        fp64_att.append(st["fp64_steps"] if opts.precision != bsa.PREC_FP64 else st["steps"])
        fallback.append(st["fallback_rays"])
    n_attr = len(integ_ms)

    # HDR colour of the last attribution frame (an HBM-bound pass: per ray it reads the
    # class byte and four doubles and writes four halves)
    k_last = frames[:: max(1, len(frames) // 24)][-1]
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tr.shade(frame_views[k_last], orb.out)
    ev0.record()
    for _ in range(10):
        tr.shade(frame_views[k_last], orb.out)
    ev1.record()
    torch.cuda.synchronize()
    shade_ms = ev0.elapsed_time(ev1) / 10
    # algorithmic bytes: every ray reads its class byte and writes 8 B; escaped rays read
    # (theta, phi) and disk rays (r_hit, g), 16 B each
    cls = orb.out["cls"] & 7
    n_esc, n_disk = int((cls == bsa.ESCAPED).sum().item()), int((cls == bsa.DISK).sum().item())
    shade_bytes = rays * (1 + 8) + 16 * (n_esc + n_disk)
    shade = {"kernel": "k_shade", "kernel_ms": round(shade_ms, 4), "bound": "hbm",
             "achieved": round(shade_bytes / (shade_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(shade_bytes / (shade_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "bytes_per_ray": round(shade_bytes / rays, 2)}

    # roofline of the dominant kernel group (HIP-event times of the attribution pass)
    avg_integ = sum(integ_ms) / n_attr
    avg_sur = sum(sur_ms) / n_attr
    att = sum(attempts) / n_attr
    att64 = min(att, sum(fp64_att) / n_attr)
    ops = att * OPS_PER_ATTEMPT
    ideal_s = (att64 * OPS_PER_ATTEMPT / FP64_VECTOR_PEAK_TFLOPS + (att - att64) * OPS_PER_ATTEMPT
               / FP32_VECTOR_PEAK_TFLOPS) / 1e12
    integ_peak = ops / ideal_s / 1e12 if ideal_s > 0 else FP64_VECTOR_PEAK_TFLOPS
    integ_ach = ops / (avg_integ * 1e-3) / 1e12 if avg_integ > 0 else 0.0
    # the fallback queues of the closed form and the surrogate run the one-shot kernel
    kname = "k_trace_oneshot" if args.path in ("closed", "surrogate") else ["k_trace_persistent",
                                                                            "k_trace_oneshot"][st["mode"]]
    wkey = f"{args.workload}/{args.path}" if args.path in ("surrogate", "closed") else args.workload
    tr_integ, src_integ = profiled_traffic(wkey, args.height, kname)
    roof_integ = {"bound": "valu", "achieved": round(integ_ach, 3), "peak": round(integ_peak, 1), "unit": "TFLOP/s",
                  "frac": round(integ_ach / integ_peak, 4), "traffic": tr_integ, "traffic_source": src_integ,
                  "kernel": kname,
                  "kernel_ms": round(avg_integ, 3), "dp5_attempts": int(att), "fp64_attempts": int(att64),
                  "ops_per_attempt": OPS_PER_ATTEMPT,
                  "peak_is": "FP64 78.6 / FP32 157.3 TFLOP/s blended by the FP64 share of attempts"}
    roof = roof_integ
    roof_mlp = None
    net_desc = None
    if args.path == "surrogate":
        hidden = 128 if tr_net == bsa.NET_FAST else 256
        flop_ray = mlp_flop_per_ray(hidden)
        mlp_ach = rays * flop_ray / (avg_sur * 1e-3) / 1e12
        kmlp = "k_mlp_fast" if hidden == 128 else "k_mlp_fused"
        net_desc = f"5-{hidden}x4-8 " + ("hard SiLU" if hidden == 128 else "SiLU")
        tr_mlp, src_mlp = profiled_traffic(wkey, args.height, kmlp)
        roof_mlp = {"bound": "mfma", "achieved": round(mlp_ach, 2), "peak": BF16_MFMA_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(mlp_ach / BF16_MFMA_PEAK_TFLOPS, 4), "traffic": tr_mlp,
                    "traffic_source": src_mlp,
                    "kernel": f"{kmlp} (features, network, gate)", "kernel_ms": round(avg_sur, 3),
                    "flop_per_ray": flop_ray,
                    "network": net_desc}
        if avg_sur > avg_integ:
            roof, roof_mlp = roof_mlp, roof_integ
    if args.path == "closed":
        # the closed-form kernels: operations per ray of the algorithm (CLOSED_FORM_OPS_PER_RAY)
        # against the vector peak of the precision they run in (FP32 157.3 TFLOP/s for the FP32
        # pass, FP64 78.6); the integrator over the rays they declined beside them
        cprec = args.closed_precision
        cops = CLOSED_FORM_OPS_PER_RAY[cprec]
        cpeak = FP32_VECTOR_PEAK_TFLOPS if cprec == "fp32" else FP64_VECTOR_PEAK_TFLOPS
        cf_ach = rays * cops / (avg_sur * 1e-3) / 1e12 if avg_sur > 0 else 0.0
        kcf = ("k_closed_form<kEvery> FP32 pass + k_closed_form<kUnsure> FP64 re-solve" if cprec == "fp32"
               else "k_closed_form<kEvery> FP64")
        # (the profile covers the FP32 pass and its FP64 re-solve)
        tr_cf, src_cf = profiled_traffic(wkey, args.height, "k_closed_form") if cprec == "fp32" else (None, None)
        roof = {"bound": "valu", "achieved": round(cf_ach, 3), "peak": cpeak, "unit": "TFLOP/s",
                "frac": round(cf_ach / cpeak, 4), "traffic": tr_cf, "traffic_source": src_cf,
                "kernel": kcf, "kernel_ms": round(avg_sur, 3), "ops_per_ray": cops,
                "ops_per_ray_is": f"mean operations of the closed form ({cprec}) per ray of the orbit frames "
                                  "(+ - * / sqrt once, FMA twice), counted by tests/closed_form.py"}
        roof_mlp = roof_integ

    desc = {"orbit": f"120-frame orbit r=40 M, i 20->85 deg, a=0.998, {width}x{args.height} Bardeen plane "
                     f"+-1.5 mean shadow radii wide, disk r_ISCO..20 M",
            "c2": f"single frame a=0.998, r=50 M, i=85 deg, {width}x{args.height} Bardeen plane 30 M wide, "
                  f"disk r_ISCO..40 M",
            "c3": f"64 seeded views (seed 0x5EED0003) of uneven size, spin and camera radius in one launch, "
                  f"{rays} rays",
            "infall": f"camera falling from rest at infinity, inside the a=0.9 horizon at r=1.2 M, theta 60 deg, "
                      f"{width}x{args.height} pinhole 120 deg (Kerr-Schild phi, far-side rule), disk r_ISCO..20 M"}[
        args.workload]
    rec = {
        "metric": METRIC if args.workload == "orbit" else f"Mrays/s, {args.workload} workload, {args.path} path",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / wall_frames, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "f64", "mixed": "f32+f64", "surrogate": "bf16+f32+f64",
                  "closed": "f32+f64" if args.closed_precision == "fp32" else "f64"}[args.path],
        "data": "synthetic: analytic Bardeen image planes, no input files; surrogate weights trained by "
                "black_simulation_amd.train_surrogate on integrator-labelled rays",
        "config": {"workload": f"{args.path}: {desc}", "path": args.path, "rays_per_step": rays,
                   "network": net_desc,
                   "fallback": (f"{'FP64' if opts.precision == bsa.PREC_FP64 else 'mixed FP32/FP64'} DP5 rtol "
                                f"{opts.rtol if opts.precision == bsa.PREC_FP64 else opts.rtol32:g}")
                   if args.path in ("surrogate", "closed") else None,
                   "gate": {"prob": opts.gate_prob, "ring": opts.gate_ring,
                            "escaped_and_disk_verified_by_closed_form": bool(opts.closed_form)}
                   if args.path == "surrogate" else None,
                   "streams": args.streams, "lean": args.lean,
                   "frame_ms_p50": round(percentile(lat, 0.5), 3), "frame_ms_p99": round(percentile(lat, 0.99), 3),
                   "plain_fp64": {"value": round(value_p, 3), "frame_ms_p50": round(percentile(lat_p, 0.5), 3),
                                  "frame_ms_p99": round(percentile(lat_p, 0.99), 3)},
                   "speedup_vs_plain_fp64": round(value / value_p, 3),
                   "host_outputs_value": round(host_value, 3) if host_value else None,
                   "host_outputs_pageable_value": round(host_value_pageable, 3) if host_value_pageable else None,
                   "host_outputs_f32_value": round(host_value_f32, 3) if host_value_f32 else None,
                   "fallback_fraction": round(sum(fallback) / n_attr / rays, 4)
                   if args.path in ("surrogate", "closed") else None,
                   "accuracy_vs_fp64": acc,
                   "closed_form": ({"fp32": "FP32 elliptic integrals, FP64 roots and the rays FP32 leaves unsure",
                                    "fp64": "FP64 throughout"}[args.closed_precision]
                                   if args.path == "closed" else None),
                   "other_paths": alt_acc or None,
                   "shade": shade,
                   "launch": {"mode": ["persistent", "one-shot"][st["mode"]], "block": st["block"],
                              "grid": st["grid"], "waves_per_cu": st["waves_per_cu"],
                              "steps_per_dequeue": st["steps_per_dequeue"], "waves_per_simd": st["waves_per_simd"]}},
        "roofline": roof,
        "roofline_other": roof_mlp,
        "cpu_baseline": None,
    }
    orb.close()
    if ref is not orb:
        ref.close()
    if not args.no_cpu and args.workload == "orbit":
        rec["c1"] = c1_timing(tr, torch)
    tr.close()
    if not args.no_cpu:
        spread = sorted(set(frames))
        # whole frames when one is too short a sample: spread over the orbit, or the one
        # frame of the c2 workload again
        more = ([frame_views[k][0] for k in spread[:: max(1, len(spread) // 8)]] if args.workload == "orbit" else
                [frame_views[0][0]] * 8 if args.workload == "c2" else [])
        rec["cpu_baseline"] = cpu_baseline(frame_views[frames[len(frames) // 2]][0], more=more)
        if args.workload == "infall":
            rec["cpu_reference"] = infall_cpu_reference(frame_views[frames[len(frames) // 2]][0])
        if args.workload in ("orbit", "c2"):
            rec["cpu_reference"] = bardeen_cpu_reference(frame_views[frames[len(frames) // 2]][0])
    print(json.dumps(rec), flush=True)


def eht_cpu_baseline(view, flow, seconds_target: float = 15.0) -> dict:
    """The oracle's EHT trace (FP64 DP5 + transfer) on this host's cores, one pixel per
    call from a thread pool (ctypes releases the GIL), over every stride-th pixel."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    cores = os.cpu_count() or 1
    ov = O.eht_view(view.a, view.inclination, view.half_width, view.size, view.frequency, view.mass_solar)
    of = O.flow(flow.n0, flow.T0, flow.beta, flow.keplerian, flow.height, flow.sin_theta_b)
    opts = O.make_opts()
    n = view.size * view.size
    px = view.pixel()

    def one(q):
        row, col = divmod(q, view.size)
        return O.eht_trace(ov, of, opts, -view.half_width + (col + 0.5) * px, view.half_width - (row + 0.5) * px)

    def run(stride):
        idx = list(range(stride // 2, n, stride))
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(one, idx, chunksize=16))
        return len(idx), time.perf_counter() - t0

    k, dt = run(max(1, n // (64 * cores)))
    stride = max(1, int(n / max(1.0, k / dt * seconds_target)))
    k, dt = run(stride)
    return {"value": round(k / dt / 1e6, 5), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th pixel of the {view.size}x{view.size} image ({k} rays, {dt:.1f} s, "
                      f"FP64 DP5 rtol 1e-10 + transfer, {cores} threads)"}


def eht_cpu_reference(view, flow, seconds_target: float = 10.0):
    """The reference's own eht::traceRay with the light (src/eht.cpp:260-319: RK4 with its
    step rule), compiled from its sources into oracle/_ref/libkerr_ref.so, on the same
    pixels from a thread pool as eht::render does (src/eht.cpp:321-347); None when that
    library was not built."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    if not os.path.exists(O.REF_LIB_PATH):
        return None
    lib = C.CDLL(O.REF_LIB_PATH)
    dp = C.POINTER(C.c_double)
    lib.ref_eht_trace_light.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp, C.c_double, C.c_double, dp]
    lib.ref_eht_trace_light.restype = None
    fl = (C.c_double * 6)(flow.n0, flow.T0, flow.beta, flow.keplerian, flow.height, flow.sin_theta_b)
    cores = os.cpu_count() or 1
    n = view.size * view.size
    px = view.pixel()

    def one(q):
        out = (C.c_double * 3)()
        row, col = divmod(q, view.size)
        lib.ref_eht_trace_light(view.a, view.inclination, view.frequency, view.mass_solar, fl,
                                -view.half_width + (col + 0.5) * px, view.half_width - (row + 0.5) * px, out)
        return out[0]

    def run(stride):
        idx = list(range(stride // 2, n, stride))
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(one, idx, chunksize=16))
        return len(idx), time.perf_counter() - t0

    k, dt = run(max(1, n // (64 * cores)))
    stride = max(1, int(n / max(1.0, k / dt * seconds_target)))
    k, dt = run(stride)
    return {"value": round(k / dt / 1e6, 5), "unit": "Mrays/s", "cores": cores, "kind": "reference",
            "sample": f"every {stride}th pixel of the {view.size}x{view.size} image ({k} rays, {dt:.1f} s): the "
                      f"reference's eht::traceRay with the light (its RK4), {cores} threads"}


def eht_bench(args, torch):
    """--workload eht: M87* at 230 GHz (a = 0.9 seen from 163 deg, 32 M field, the flow's
    density calibrated to 0.5 Jy), size x size pixels per step, FP64 throughout."""
    from black_simulation_amd import eht

    tr = __import__("black_simulation_amd").Tracer(0)
    m87 = eht.TARGETS[0]
    view = eht.target_view(m87, a=0.9, half_width=16.0, size=args.size)
    flow = eht.Flow()
    theta_g = eht.angular_scale(m87.mass_solar, m87.distance_pc)
    cal_ms = None
    if args.lean:  # for profilers: only the timed renders (n0 as calibrated for this view)
        flow.n0 = 1.7544e5
    else:
        eht.calibrate(view, flow, m87.flux, theta_g, tracer=tr)  # warm
        t0 = time.perf_counter()
        flow.n0 = eht.calibrate(view, flow, m87.flux, theta_g, tracer=tr)
        cal_ms = round((time.perf_counter() - t0) * 1e3, 2)
    rays = args.size * args.size
    for _ in range(args.warmup):
        eht.render(view, flow, tracer=tr)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        img = eht.render(view, flow, tracer=tr)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    value = rays * args.steps / wall / 1e6
    jy = eht.flux(img, theta_g, tracer=tr)
    if args.lean:
        print(json.dumps({"kernel_ms": kernel_ms, "value": value}), flush=True)
        tr.close()
        return
    det = eht.render(view, flow, details=True, tracer=tr)
    torch.cuda.synchronize()
    attempts = int(det.steps.to(torch.int64).sum().item())
    ach = attempts * OPS_PER_ATTEMPT / (kernel_ms * 1e-3) / 1e12
    rec = {
        "metric": "Mrays/s EHT image at 230 GHz (M87*, a=0.9), FP64 geodesics + synchrotron transfer",
        "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic: analytic Bardeen image plane and flow model",
        "config": {"workload": f"eht: M87* a=0.9 i=163 deg, {args.size}x{args.size} pixels over 32 M, 230 GHz, "
                               f"n0 calibrated to 0.5 Jy on 48x48", "rays_per_step": rays,
                   "calibrate_ms": cal_ms, "n0": flow.n0, "flux_jy": round(jy, 4),
                   "captured_fraction": round(float(det.shadow.float().mean().item()), 4),
                   "mean_attempts_per_ray": round(attempts / rays, 1)},
        "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / FP64_VECTOR_PEAK_TFLOPS, 4), "traffic": profiled_traffic("eht", args.size, "k_eht")[0],
                     "traffic_source": profiled_traffic("eht", args.size, "k_eht")[1], "kernel": "k_eht",
                     "kernel_ms": round(kernel_ms, 3), "dp5_attempts": attempts, "ops_per_attempt": OPS_PER_ATTEMPT,
                     "note": "counts the DP5 operations only; the transfer's exp/pow/Bessel work comes on top"},
        "cpu_baseline": None,
    }
    tr.close()
    if not args.no_cpu:
        rec["cpu_baseline"] = eht_cpu_baseline(view, flow)
        rec["cpu_reference"] = eht_cpu_reference(view, flow)
    print(json.dumps(rec), flush=True)


def debris_bench(args, torch):
    """--workload debris: 2^22 particles (a = 0.9) launched from the local ZAMOs at r in
    [2.5, 40] M with speeds up to 0.75 c, one step = one debris update of dt = 10 M of
    Kerr-Schild time for every moving particle, FP64."""
    import numpy as np

    import black_simulation_amd as bsa
    from black_simulation_amd.particles import COLUMNS, Particles

    a, n, dt = 0.9, args.particles, 10.0
    rng = np.random.default_rng(0xDEB2)
    r = rng.uniform(2.5, 40.0, n)
    th = np.arccos(rng.uniform(-0.95, 0.95, n))
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    v = rng.normal(size=(n, 3))
    v *= (rng.uniform(0.0, 0.75, n) / np.linalg.norm(v, axis=1))[:, None]
    # the launch closed form, vectorised (kerr::launchFromZamo with mu = 1, as kt_particle_from_zamo)
    sn, cs = np.sin(th), np.cos(th)
    sig, dl = r * r + a * a * cs * cs, r * r - 2 * r + a * a
    A = (r * r + a * a) ** 2 - a * a * dl * sn * sn
    g = 1.0 / np.sqrt(1.0 - np.minimum((v * v).sum(1), 1 - 1e-12))
    L = g * v[:, 2] * np.sqrt(A / sig) * sn
    E = g * np.sqrt(sig * dl / A) + 2 * a * r / A * L
    vth = g * v[:, 1] * np.sqrt(sig)
    cols = {"r": r, "mu": cs, "phi": ph, "vr": g * v[:, 0] * np.sqrt(sig * dl), "vmu": -sn * vth, "E": E, "L": L,
            "Q": vth * vth + cs * cs * (a * a * (1 - E * E) + L * L / np.maximum(sn * sn, 1e-18))}
    tr = bsa.Tracer(0)
    p = Particles(cols)
    start = {c: getattr(p, c).clone() for c in COLUMNS}
    fate0 = p.fate.clone()

    def reset():
        for c in COLUMNS:
            getattr(p, c).copy_(start[c])
        p.fate.copy_(fate0)
        p.h.zero_()

    for _ in range(args.warmup):
        p.advance(tr, a, dt)
    reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p.advance(tr, a, dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = n * args.steps / wall / 1e6
    # attribution: the same updates again, each between HIP events on the stream, with the
    # attempts each update took (the steps column)
    reset()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    attempts = []
    for e0, e1 in ev:
        e0.record(stream)
        p.advance(tr, a, dt)
        e1.record(stream)
        attempts.append(p.steps.to(torch.int64).sum())
    torch.cuda.synchronize()
    kernel_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / args.steps
    attempts = int(sum(int(x.item()) for x in attempts))
    moving = int((p.fate == 0).sum().item())
    ach = attempts * OPS_PER_PARTICLE_ATTEMPT / (kernel_ms * 1e-3 * args.steps) / 1e12
    rec = {
        "metric": "Mparticle-updates/s, debris workload (a=0.9, dt=10 M per update), FP64",
        "value": round(value, 3), "unit": "Mparticles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic: seeded debris launched from local ZAMOs",
        "config": {"workload": f"debris: {n} particles, a=0.9, r 2.5..40 M, |v| < 0.75 c, dt=10 M per update",
                   "particles": n, "moving_after": moving,
                   "mean_attempts_per_update": round(attempts / max(1, n * args.steps), 2)},
        "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / FP64_VECTOR_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic("debris", args.particles, "k_advance")[0],
                     "traffic_source": profiled_traffic("debris", args.particles, "k_advance")[1], "kernel": "k_advance",
                     "kernel_ms": round(kernel_ms, 3), "dp5_attempts": attempts,
                     "ops_per_attempt": OPS_PER_PARTICLE_ATTEMPT,
                     "note": "attempts of all timed updates over their HIP-event time (persistent k_advance)"},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        sample = {c: start[c][: 1 << 16].cpu().numpy() for c in COLUMNS}
        rec["cpu_baseline"] = debris_cpu_baseline(a, sample, dt)
        rec["cpu_reference"] = debris_cpu_reference(a, sample, dt)
    tr.close()
    print(json.dumps(rec), flush=True)


def debris_cpu_baseline(a, cols, dt, seconds_target: float = 10.0) -> dict:
    """The oracle's particle advance (same DP5) on the host cores, one particle per call
    from a thread pool, over the first particles of the workload."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    cores = os.cpu_count() or 1
    opts = O.make_opts()
    rs = O.lib().ko_horizon_radius(a) + 0.02

    def one(i):
        st = [cols[c][i] for c in ("r", "mu", "phi", "vr", "vmu")]
        return O.particle_advance(a, cols["E"][i], cols["L"][i], cols["Q"][i], st, dt, rs, 1e4, opts)[0]

    def run(k):
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(one, range(k), chunksize=64))
        return time.perf_counter() - t0

    k = min(len(cols["r"]), 4096)
    dt0 = run(k)
    k = int(min(len(cols["r"]), max(k, k / dt0 * seconds_target)))
    t = run(k)
    return {"value": round(k / t / 1e6, 5), "unit": "Mparticles/s", "cores": cores, "kind": "port",
            "sample": f"the first {k} particles of the workload, one update of dt={dt} M each ({t:.1f} s, {cores} threads)"}


def debris_cpu_reference(a, cols, dt, seconds_target: float = 10.0):
    """The reference's own trajectory::predict (src/trajectory.cpp:121-172: RK4 in
    coordinate time) over dt for each of the workload's first particles, from a thread
    pool; None when oracle/_ref was not built."""
    import ctypes as C

    lib = _ref_lib()
    if lib is None:
        return None
    dp = C.POINTER(C.c_double)
    lib.ref_predict.argtypes = [C.c_double, dp, dp, C.c_double, C.c_double, dp]
    lib.ref_predict.restype = None

    def one(i):
        mu = min(1.0, max(-1.0, float(cols["mu"][i])))
        th = math.acos(mu)
        sn = max(math.sin(th), 1e-300)
        consts = (C.c_double * 4)(cols["E"][i], cols["L"][i], cols["Q"][i], 1.0)
        state = (C.c_double * 6)(0.0, cols["r"][i], th, cols["phi"][i], cols["vr"][i], -cols["vmu"][i] / sn)
        out = (C.c_double * 9)()
        lib.ref_predict(a, consts, state, dt, 1e4, out)
        return out[0]

    k, t, stride, cores = _timed_pool(one, len(cols["r"]), seconds_target)
    return {"value": round(k / t / 1e6, 5), "unit": "Mparticles/s", "cores": cores, "kind": "reference",
            "sample": f"every {stride}th of the first {len(cols['r'])} particles ({k}), one update of dt={dt} M each "
                      f"({t:.1f} s): the reference's trajectory::predict (its RK4), {cores} threads"}


if __name__ == "__main__":
    main()
