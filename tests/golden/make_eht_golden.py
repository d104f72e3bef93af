"""Writes tests/golden/eht_vectors.npz from the reference's own EHT code (src/eht.cpp).

Runs only where /root/reference exists: oracle/Makefile's `ref` target builds
oracle/_ref/libkerr_ref.so from the reference's sources read in place plus
oracle/ref_shim.cpp.  The fixture holds inputs and the reference's outputs only:

  - besselK0/K1/K2 on both sides of their branch point;
  - synchrotronEmissivity and planck on seeded physical inputs (including the cut-offs);
  - plasmaAt and flowVelocity on seeded (r, theta, a) inside and outside the ISCO;
  - traceRay with the light for an 8 x 8 grid of two views (RK4 with its own step rule:
    a loose check of the transfer, not a bit-level one);
  - calibrate's n0 for one small view.

    python tests/golden/make_eht_golden.py
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.join(HERE, "eht_vectors.npz")
FLOW_DEFAULT = [1e5, 4e11, 10.0, 0.9, 1.0, 0.8165]
FLOW_THIN = [3e5, 2e11, 5.0, 0.7, 0.3, 0.6]


def load_ref():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    L = C.CDLL(os.path.join(ORACLE, "_ref", "libkerr_ref.so"))
    d, i, P = C.c_double, C.c_int, C.POINTER(C.c_double)
    for name, res, args in [
        ("ref_bessel_k", d, [i, d]),
        ("ref_synchrotron_emissivity", d, [d, d, d, d, d]),
        ("ref_planck_nu", d, [d, d]),
        ("ref_plasma_at", None, [P, d, d, d, P]),
        ("ref_flow_velocity", None, [P, d, d, d, P]),
        ("ref_eht_trace_light", None, [d, d, d, d, P, d, d, P]),
        ("ref_eht_calibrate", d, [d, d, d, d, d, P, d, d]),
        ("ref_horizon_radius", d, [d]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def main():
    L = load_ref()
    rng = np.random.default_rng(0xE47)
    out = {}

    x = np.concatenate([np.logspace(-2, math.log10(2.0), 20), np.logspace(math.log10(2.0001), 2, 20)])
    out["bessel_x"] = x
    out["bessel_k"] = np.array([[L.ref_bessel_k(w, v) for v in x] for w in range(3)])

    n = 300
    em = np.stack([10 ** rng.uniform(9, 13, n), 10 ** rng.uniform(2, 8, n), 10 ** rng.uniform(-1.6, 2, n),
                   10 ** rng.uniform(-2, 3, n), rng.uniform(0.0, 1.0, n)], axis=1)
    em[:10, 2] = rng.uniform(0.0, 0.05, 10)  # below the thetaE cut
    out["emis_in"] = em
    out["emis"] = np.array([L.ref_synchrotron_emissivity(*row) for row in em])
    pl = np.stack([10 ** rng.uniform(8, 14, n), 10 ** rng.uniform(3, 12, n)], axis=1)
    out["planck_in"] = pl
    out["planck"] = np.array([L.ref_planck_nu(*row) for row in pl])

    rows = []
    for _ in range(n):
        a = rng.uniform(0.0, 0.998)
        rp = L.ref_horizon_radius(a)
        r = rp * 1.02 + (60.0 - rp) * rng.uniform() ** 3
        rows.append([a, r, rng.uniform(0.02, math.pi - 0.02)])
    geo = np.array(rows)
    out["flow_geo"] = geo
    for tag, flow in (("default", FLOW_DEFAULT), ("thin", FLOW_THIN)):
        f = np.array(flow)
        pa, uu = [], []
        for a, r, th in geo:
            p = np.zeros(3)
            u = np.zeros(4)
            L.ref_plasma_at(ptr(f), r, th, a, ptr(p))
            L.ref_flow_velocity(ptr(f), a, r, th, ptr(u))
            pa.append(p)
            uu.append(u)
        out[f"flow_{tag}"] = f
        out[f"plasma_{tag}"] = np.array(pa)
        out[f"velocity_{tag}"] = np.array(uu)

    # two views: M87*-like (a 0.9 seen from 163 deg) and a face-on Sgr A*-like one
    views = np.array([[0.9, math.radians(163.0), 16.0, 230e9, 6.5e9],
                      [0.5, math.radians(30.0), 16.0, 230e9, 4.3e6]])
    flows = np.array([[6e4, 4e11, 10.0, 0.9, 1.0, 0.8165], [3e7, 4e11, 10.0, 0.9, 1.0, 0.8165]])
    grid = (np.arange(8) + 0.5) / 8 * 2 - 1
    t0 = time.time()
    traced = []
    for (a, inc, hw, nu, mass), f in zip(views, flows):
        res = []
        for b in grid[::-1] * hw:
            for al in grid * hw:
                o = np.zeros(3)
                L.ref_eht_trace_light(a, inc, nu, mass, ptr(f), al, b, ptr(o))
                res.append([al, b, *o])
        traced.append(res)
    out["trace_views"] = views
    out["trace_flows"] = flows
    out["trace"] = np.array(traced)  # [view][ray] = alpha, beta, intensity, captured, minRadius
    print(f"traced {out['trace'].shape} in {time.time() - t0:.1f} s")

    t0 = time.time()
    cal_view = np.array([0.9, math.radians(163.0), 16.0, 230e9, 6.5e9])
    theta_g = 6.5e9 * 1.476625e5 / (16.8e6 * 3.0856775814913673e18) / (math.pi / 180 / 3600 / 1e6)
    cal_flow = np.array(FLOW_DEFAULT)
    out["calibrate_in"] = np.concatenate([cal_view, [0.5, theta_g]])
    out["calibrate_flow"] = cal_flow
    out["calibrate_n0"] = np.array(L.ref_eht_calibrate(*cal_view, ptr(cal_flow), 0.5, theta_g))
    print(f"calibrate n0 = {float(out['calibrate_n0']):.4e} in {time.time() - t0:.1f} s")

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
