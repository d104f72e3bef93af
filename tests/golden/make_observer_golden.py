"""Writes tests/golden/observer_vectors.npz from the reference's own code for observers other
than the static ZAMO: boosted ZAMOs (Camera::frame's Hover with drift, src/camera.cpp:23-36)
and free-falling cameras, also inside the horizon (src/trajectory.cpp:234-241).

Runs only where /root/reference exists: oracle/Makefile's `ref` target builds
oracle/_ref/libkerr_ref.so from the reference's sources read in place plus
oracle/ref_shim.cpp.  The fixture holds inputs and the reference's outputs only:

  - kerr::fourVelocity for massive geodesic states, outside and inside the horizon;
  - kerr::derivatives (Kerr-Schild dphi) for photon states inside and near the horizon;
  - trajectory::traceRay from boosted and falling observers' frames (RK4 with its own
    step rule: class, the sky chord and the launch constants E, L/E);
  - kerr::observerFrame of those observers (the frames themselves).

    python tests/golden/make_observer_golden.py
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.join(HERE, "observer_vectors.npz")


def load_ref():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    L = C.CDLL(os.path.join(ORACLE, "_ref", "libkerr_ref.so"))
    d, i, P = C.c_double, C.c_int, C.POINTER(C.c_double)
    for name, res, args in [
        ("ref_horizon_radius", d, [d]),
        ("ref_isco_radius", d, [d, i]),
        ("ref_zamo_frame", None, [d, d, d, P]),
        ("ref_derivatives", None, [d, P, P, P]),
        ("ref_radial_potential", d, [d, P, d]),
        ("ref_polar_potential", d, [d, P, d]),
        ("ref_four_velocity", None, [d, P, P, P]),
        ("ref_trajectory_trace_u", None, [d, d, d, d, P, P, i, d, d, d, P]),
        ("ref_observer_frame", None, [d, d, d, P, P]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def ptr(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))


def arr(x):
    return np.ascontiguousarray(np.asarray(x, np.float64))


def main():
    L = load_ref()
    rng = np.random.default_rng(0x0B5E)
    out = {}

    # kerr::fourVelocity of massive geodesics (mu = 1): [a, E, L, Q, r, theta, sign_r, sign_th] -> u
    rows, us = [], []
    while len(rows) < 200:
        a = rng.uniform(0.0, 0.998)
        rp = L.ref_horizon_radius(a)
        rm = 2.0 - rp
        inside = len(rows) % 3 == 0
        r = rng.uniform(rm + 0.05 * (rp - rm), rp * 0.999) if inside else rng.uniform(rp * 1.01, 30.0)
        th = rng.uniform(0.1, math.pi - 0.1)
        E, Lz, Q = rng.uniform(0.6, 1.4), rng.uniform(-4, 4), rng.uniform(0, 12)
        c = arr([E, Lz, Q, 1.0])
        R = L.ref_radial_potential(a, ptr(c), r)
        T = L.ref_polar_potential(a, ptr(c), th)
        if R <= 0 or T <= 0:
            continue
        sr, st = rng.choice([-1, 1]), rng.choice([-1, 1])
        s = arr([0.0, r, th, 0.0, sr * math.sqrt(R), st * math.sqrt(T)])
        u = np.zeros(4)
        L.ref_four_velocity(a, ptr(c), ptr(s), u.ctypes.data_as(C.POINTER(C.c_double)))
        rows.append([a, E, Lz, Q, r, th, sr, st])
        us.append(u)
    out["fourvel_in"] = np.array(rows)
    out["fourvel"] = np.array(us)

    # Kerr-Schild dphi of photons inside and near the horizon: [a, E, L, Q, r, theta, vr, vth] -> derivs
    rows, ds = [], []
    while len(rows) < 300:
        a = rng.uniform(0.05, 0.998)
        rp = L.ref_horizon_radius(a)
        rm = 2.0 - rp
        k = len(rows) % 3
        r = (rng.uniform(rm + 0.05 * (rp - rm), rp * 0.999) if k == 0 else
             rp * (1.0 + 10 ** rng.uniform(-6, -2)) if k == 1 else rng.uniform(rp * 1.01, 12.0))
        th = rng.uniform(0.1, math.pi - 0.1)
        E, Lz, Q = 1.0, rng.uniform(-5, 5), rng.uniform(0, 30)
        c = arr([E, Lz, Q, 0.0])
        R = L.ref_radial_potential(a, ptr(c), r)
        T = L.ref_polar_potential(a, ptr(c), th)
        if R <= 0 or T <= 0:
            continue
        vr = rng.choice([-1, 1]) * math.sqrt(R)
        vth = rng.choice([-1, 1]) * math.sqrt(T)
        s = arr([0.0, r, th, 0.3, vr, vth])
        d = np.zeros(6)
        L.ref_derivatives(a, ptr(c), ptr(s), d.ctypes.data_as(C.POINTER(C.c_double)))
        rows.append([a, E, Lz, Q, r, th, vr, vth])
        ds.append(d)
    out["ksrhs_in"] = np.array(rows)
    out["ksrhs"] = np.array(ds)

    # trajectory::traceRay from observers' frames: [a, r, theta, u0..u3, look0..2, disk, r_in, r_out, esc]
    cases, res = [], []
    looks = rng.normal(size=(16, 3))
    looks /= np.linalg.norm(looks, axis=1, keepdims=True)
    looks[0] = [1.0, 0.0, 0.0]   # radially outward: the photon comes from larger r
    looks[1] = [-1.0, 0.0, 0.0]  # radially inward

    def zamo_boost(a, r, th, v):
        e = np.zeros(16)
        L.ref_zamo_frame(a, r, th, e.ctypes.data_as(C.POINTER(C.c_double)))
        e = e.reshape(4, 4)
        g = 1.0 / math.sqrt(1.0 - float(np.dot(v, v)))
        return g * (e[0] + v[0] * e[1] + v[1] * e[2] + v[2] * e[3])

    def falling(a, r, th, E, Lz, Q, sr=-1, st=1):
        c = arr([E, Lz, Q, 1.0])
        R = L.ref_radial_potential(a, ptr(c), r)
        T = L.ref_polar_potential(a, ptr(c), th)
        s = arr([0.0, r, th, 0.0, sr * math.sqrt(max(R, 0.0)), st * math.sqrt(max(T, 0.0))])
        u = np.zeros(4)
        L.ref_four_velocity(a, ptr(c), ptr(s), u.ctypes.data_as(C.POINTER(C.c_double)))
        return u

    observers = []
    a = 0.9
    isco = L.ref_isco_radius(a, 1)
    for v in ([0.0, 0.0, 0.6], [-0.5, 0.0, 0.0], [0.0, 0.3, 0.0], [0.3, -0.2, 0.4]):
        observers.append((a, 10.0, 1.2, zamo_boost(a, 10.0, 1.2, np.array(v)), 1, isco, 20.0))
    rp = L.ref_horizon_radius(a)
    for r in (1.2, rp * 0.999, rp * 1.001, 1.6):  # falling from rest at infinity, across the horizon
        observers.append((a, r, 1.0, falling(a, r, 1.0, 1.0, 0.0, 0.0), 1, isco, 20.0))
    observers.append((a, 1.2, 1.3, falling(a, 1.2, 1.3, 1.0, 2.0, 3.0), 0, isco, 20.0))
    observers.append((0.0, 1.5, 1.4, falling(0.0, 1.5, 1.4, 1.0, 0.0, 0.0), 0, 6.0, 20.0))
    for a_, r_, th_, u, disk, rin, rout in observers:
        for lk in looks:
            o = np.zeros(11)
            L.ref_trajectory_trace_u(a_, r_, th_, 0.0, ptr(u), ptr(lk), disk, rin, rout, 1000.0,
                                     o.ctypes.data_as(C.POINTER(C.c_double)))
            cases.append([a_, r_, th_, *u, *lk, disk, rin, rout, 1000.0])
            res.append(o)
    frames_in, frames = [], []
    for a_, r_, th_, u, disk, rin, rout in observers:
        e = np.zeros(16)
        L.ref_observer_frame(a_, r_, th_, ptr(u), e.ctypes.data_as(C.POINTER(C.c_double)))
        frames_in.append([a_, r_, th_, *u])
        frames.append(e)
    out["frame_in"] = np.array(frames_in)
    out["frame_out"] = np.array(frames)
    out["trace_cases"] = np.array(cases)
    out["trace_out"] = np.array(res)
    ends = np.array(res)[:, 0].astype(int)
    print("traceRay ends:", {k: int((ends == k).sum()) for k in sorted(set(ends))})

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
