"""Writes tests/golden/particle_vectors.npz from the reference's own code for massive
particles (mu = 1): kerr::launchFromZamo and trajectory::predict (coordinate-time RK4 with
kerr::suggestedTimeStep, src/trajectory.cpp:121-172).

Runs only where /root/reference exists: oracle/Makefile's `ref` target builds
oracle/_ref/libkerr_ref.so from the reference's sources read in place plus
oracle/ref_shim.cpp.  The fixture holds inputs and the reference's outputs only.

    python tests/golden/make_particle_golden.py
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
OUT = os.path.join(HERE, "particle_vectors.npz")


def main():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    L = C.CDLL(os.path.join(ORACLE, "_ref", "libkerr_ref.so"))
    d, P = C.c_double, C.POINTER(C.c_double)
    L.ref_horizon_radius.restype = d
    L.ref_horizon_radius.argtypes = [d]
    L.ref_particle_from_zamo.restype = None
    L.ref_particle_from_zamo.argtypes = [d, d, d, P, P]
    L.ref_predict.restype = None
    L.ref_predict.argtypes = [d, P, P, d, d, P]

    def ptr(x):
        return x.ctypes.data_as(P)

    rng = np.random.default_rng(0xDEB1)
    out = {}
    rows, res = [], []
    for _ in range(200):
        a = rng.uniform(0.0, 0.998)
        r = L.ref_horizon_radius(a) * rng.uniform(1.05, 1.5) if len(rows) % 4 == 0 else rng.uniform(3.0, 40.0)
        th = rng.uniform(0.05, math.pi - 0.05)
        v = rng.normal(size=3)
        v *= rng.uniform(0.0, 0.95) / np.linalg.norm(v)
        o = np.zeros(5)
        L.ref_particle_from_zamo(a, r, th, ptr(np.ascontiguousarray(v)), ptr(o))
        rows.append([a, r, th, *v])
        res.append(o)
    out["zamo_in"] = np.array(rows)
    out["zamo_out"] = np.array(res)

    # orbits: [a, r, theta, v0, v1, v2, max_time, escape] -> predict summary
    cases = [
        (0.9, 12.0, math.pi / 2, (0.0, 0.0, 0.30), 300.0, 2000.0),   # bound, eccentric, equatorial
        (0.9, 10.0, 1.2, (0.1, 0.2, 0.3), 300.0, 2000.0),            # bound, inclined
        (0.5, 10.0, 1.0, (0.0, 0.35, 0.0), 300.0, 2000.0),           # L = 0: over the poles
        (0.9, 6.0, 1.3, (-0.5, 0.0, 0.1), 300.0, 2000.0),            # plunges
        (0.7, 20.0, 1.1, (0.9, 0.0, 0.1), 3000.0, 400.0),            # escapes
        (0.0, 15.0, math.pi / 2, (0.0, 0.0, 0.25), 400.0, 2000.0),   # Schwarzschild, eccentric
        (0.998, 4.0, 1.4, (0.0, 0.05, 0.45), 200.0, 2000.0),         # near-extremal, near the hole
    ]
    rows, res = [], []
    for a, r, th, v, tmax, esc in cases:
        o = np.zeros(5)
        L.ref_particle_from_zamo(a, r, th, ptr(np.array(v, np.float64)), ptr(o))
        consts = np.array([o[0], o[1], o[2], 1.0])
        state = np.array([0.0, r, th, 0.0, o[3], o[4]])
        p = np.zeros(9)
        L.ref_predict(a, ptr(consts), ptr(state), tmax, esc, ptr(p))
        rows.append([a, r, th, *v, tmax, esc])
        res.append(p)
    out["predict_in"] = np.array(rows)
    out["predict_out"] = np.array(res)
    print("fates", [int(x[0]) for x in res])
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
