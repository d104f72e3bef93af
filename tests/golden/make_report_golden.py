"""Writes tests/golden/report_vectors.npz: the reference's own trajectory::traceRay
RayReport (src/trajectory.h:97-113, src/trajectory.cpp:214-396) for single rays from ZAMO
cameras: where each ended, and the turns, rMin and delay the inspector prints
(src/app.cpp:776-777).

Runs only where /root/reference exists: oracle/Makefile's `ref` target builds
oracle/_ref/libkerr_ref.so from the reference's sources read in place plus
oracle/ref_shim.cpp.  The fixture holds inputs and the reference's outputs only.  Cases:
the rays of the reference's testRayTrace (test_kerr.cpp:496-543), and seeded looks from
ZAMO cameras at four spins, with and without a disk.

    python tests/golden/make_report_golden.py
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
OUT = os.path.join(HERE, "report_vectors.npz")


def main():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    L = C.CDLL(os.path.join(ORACLE, "_ref", "libkerr_ref.so"))
    d, i, P = C.c_double, C.c_int, C.POINTER(C.c_double)
    L.ref_trajectory_report.restype = None
    L.ref_trajectory_report.argtypes = [d, d, d, d, P, i, d, d, d, P]
    L.ref_isco_radius.restype = d
    L.ref_isco_radius.argtypes = [d, i]

    cases = []  # a, r, theta, look x, y, z, disk, r_in, r_out, escape
    # testRayTrace: at the hole, sideways, radial at a = 0, grazing and wide at r = 1000, disk
    r_far, bc = 1000.0, 3.0 * math.sqrt(3.0)

    def look_for(b):
        sa = b * math.sqrt(1.0 - 2.0 / r_far) / r_far
        return [-math.sqrt(1.0 - sa * sa), 0.0, sa]

    cases.append((0.9, 30.0, math.pi / 2, [-1.0, 0.0, 0.0], 0, 2.0, 20.0, 250.0))
    cases.append((0.9, 30.0, math.pi / 2, [0.0, 0.0, 1.0], 0, 2.0, 20.0, 250.0))
    cases.append((0.0, 30.0, math.pi / 2, [-1.0, 0.0, 0.0], 0, 2.0, 20.0, 250.0))
    cases.append((0.0, r_far, math.pi / 2, look_for(bc * (1 + 1e-6)), 0, 2.0, 20.0, 250.0))
    cases.append((0.0, r_far, math.pi / 2, look_for(2 * bc), 0, 2.0, 20.0, 250.0))
    cases.append((0.9, 30.0, 1.35, [-0.95, 0.3122, 0.0], 1, L.ref_isco_radius(0.9, 1), 20.0, 250.0))
    # seeded looks: mostly toward the hole, so that many rays bend strongly or land on the disk
    rng = np.random.default_rng(20261204)
    for a in (0.0, 0.5, 0.9, 0.998):
        for k in range(10):
            r = float(rng.choice([12.0, 30.0, 80.0]))
            th = float(rng.uniform(0.3, 2.8))
            v = np.array([-1.0, 0.0, 0.0]) + rng.normal(0.0, 0.35, 3) * (1.5 if k % 3 == 0 else 0.6)
            v /= np.linalg.norm(v)
            disk = k % 2
            cases.append((a, r, th, v.tolist(), disk, L.ref_isco_radius(a, 1), 20.0, 250.0))
    out = np.zeros((len(cases), 9))
    for k, (a, r, th, lk, disk, rin, rout, esc) in enumerate(cases):
        lkv = np.ascontiguousarray(np.array(lk, dtype=float))
        o = np.zeros(9)
        L.ref_trajectory_report(a, r, th, 0.0, lkv.ctypes.data_as(P), disk, rin, rout, esc, o.ctypes.data_as(P))
        out[k] = o
    g = {
        "cases": np.array([[c[0], c[1], c[2], *c[3], c[4], c[5], c[6], c[7]] for c in cases]),
        "report": out,
    }
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT}: {len(cases)} rays; ends {np.bincount(out[:, 0].astype(int)).tolist()}")


if __name__ == "__main__":
    main()
