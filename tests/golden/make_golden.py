"""Writes tests/golden/ref_vectors.npz from the reference's own Kerr code.

Runs only in a container where /root/reference exists: oracle/Makefile's `ref` target
compiles the reference's src/kerr.cpp, eht.cpp, trajectory.cpp and central.cpp (read in
place) with oracle/ref_shim.cpp into oracle/_ref/libkerr_ref.so, which this script calls
through ctypes.  The fixture holds only inputs and the reference's outputs; the tests
compare the oracle (oracle/kerr_oracle.c) against them.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.join(HERE, "ref_vectors.npz")


def load_ref():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    L = C.CDLL(os.path.join(ORACLE, "_ref", "libkerr_ref.so"))
    d, i, P = C.c_double, C.c_int, C.POINTER(C.c_double)
    sig = {
        "ref_horizon_radius": (d, [d]),
        "ref_isco_radius": (d, [d, i]),
        "ref_photon_orbit_radius": (d, [d, i]),
        "ref_marginally_bound_radius": (d, [d, i]),
        "ref_circular_omega": (d, [d, d, i]),
        "ref_azimuth_shift": (d, [d, d]),
        "ref_derivatives": (None, [d, P, P, P]),
        "ref_radial_potential": (d, [d, P, d]),
        "ref_polar_potential": (d, [d, P, d]),
        "ref_enforce_constraints": (None, [d, P, P]),
        "ref_zamo_frame": (None, [d, d, d, P]),
        "ref_metric_tensor": (None, [d, d, d, P]),
        "ref_photon_from_zamo": (None, [d, d, d, d, P, P, P]),
        "ref_eht_trace": (None, [d, d, d, d, P]),
        "ref_critical_curve": (i, [d, d, i, P, P, i]),
        "ref_critical_radius": (d, [P, P, i, d]),
        "ref_trajectory_trace": (None, [d, d, d, d, P, i, d, d, d, P]),
        "ref_bending_angle": (d, [d, d]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def main():
    L = load_ref()
    rng = np.random.default_rng(20261203)
    g = {}

    # radii and circular orbits (kerr.cpp:157-223, 296-299)
    spins = np.array([0.0, 0.3, 0.5, 0.6, 0.9, 0.94, 0.998, 1.0])
    g["radii_spin"] = spins
    g["radii"] = np.array([[L.ref_horizon_radius(a), L.ref_isco_radius(a, 1), L.ref_isco_radius(a, 0),
                            L.ref_photon_orbit_radius(a, 1), L.ref_photon_orbit_radius(a, 0),
                            L.ref_marginally_bound_radius(a, 1)] for a in spins])
    om_in = np.array([[a, r] for a in (0.0, 0.5, 0.9, 0.998) for r in (2.0, 3.7, 6.0, 15.0, 40.0)])
    g["omega_in"] = om_in
    g["omega"] = np.array([[L.ref_circular_omega(a, r, 1), L.ref_circular_omega(a, r, 0)] for a, r in om_in])
    g["azshift_in"] = np.array([[a, r] for a in (0.5, 0.9, 0.998) for r in (3.0, 30.0, 1000.0)])
    g["azshift"] = np.array([L.ref_azimuth_shift(a, r) for a, r in g["azshift_in"]])

    # geodesic right-hand side and potentials at seeded random states (kerr.cpp:331-363)
    n = 600
    a = rng.uniform(0.0, 0.998, n)
    consts = np.stack([rng.uniform(0.5, 1.5, n), rng.uniform(-8, 8, n), rng.uniform(-2, 40, n),
                       (np.arange(n) % 3 == 0).astype(float)], axis=1)  # every third massive
    state = np.stack([np.zeros(n), rng.uniform(1.2, 60, n), rng.uniform(0.05, math.pi - 0.05, n),
                      rng.uniform(-3, 3, n), rng.uniform(-50, 50, n), rng.uniform(-6, 6, n)], axis=1)
    der = np.zeros((n, 6))
    Rp = np.zeros(n)
    Tp = np.zeros(n)
    enf = state.copy()
    for k in range(n):
        c = np.ascontiguousarray(consts[k])
        s = np.ascontiguousarray(state[k])
        out = np.zeros(6)
        L.ref_derivatives(a[k], ptr(c), ptr(s), ptr(out))
        der[k] = out
        Rp[k] = L.ref_radial_potential(a[k], ptr(c), s[1])
        Tp[k] = L.ref_polar_potential(a[k], ptr(c), s[2])
        e = np.ascontiguousarray(state[k].copy())
        L.ref_enforce_constraints(a[k], ptr(c), ptr(e))
        enf[k] = e
    g.update(rhs_a=a, rhs_consts=consts, rhs_state=state, rhs_derivs=der, rhs_R=Rp, rhs_Theta=Tp,
             enforce_state=enf)

    # metric and ZAMO frames, photons seen by a ZAMO (kerr.cpp:238-254, 424-483)
    m = 40
    fa = rng.uniform(0.0, 0.998, m)
    fr = rng.uniform(2.5, 80.0, m)
    ft = rng.uniform(0.1, math.pi - 0.1, m)
    frames = np.zeros((m, 16))
    metrics = np.zeros((m, 16))
    for k in range(m):
        e = np.zeros(16)
        L.ref_zamo_frame(fa[k], fr[k], ft[k], ptr(e))
        frames[k] = e
        gm = np.zeros(16)
        L.ref_metric_tensor(fa[k], fr[k], ft[k], ptr(gm))
        metrics[k] = gm
    g.update(frame_a=fa, frame_r=fr, frame_theta=ft, frames=frames, metrics=metrics)
    looks = rng.normal(size=(m * 5, 3))
    looks /= np.linalg.norm(looks, axis=1, keepdims=True)
    looks[::7, 2] = 1e-6  # some nearly through the axis (snap)
    looks /= np.linalg.norm(looks, axis=1, keepdims=True)
    pc = np.zeros((m * 5, 4))
    ps = np.zeros((m * 5, 6))
    for k in range(m * 5):
        j = k // 5
        lk = np.ascontiguousarray(looks[k])
        c = np.zeros(4)
        s = np.zeros(6)
        L.ref_photon_from_zamo(fa[j], fr[j], ft[j], 0.25, ptr(lk), ptr(c), ptr(s))
        pc[k] = c
        ps[k] = s
    g.update(photon_looks=looks, photon_consts=pc, photon_state=ps)

    # eht::traceRay capture flags on coarse image planes (camera at r = 1000 M)
    planes = [(0.9, math.radians(80.0), 12.0), (0.998, math.radians(85.0), 9.0), (0.0, math.radians(40.0), 8.0),
              (0.9, math.radians(17.0), 8.0)]
    cap = []
    for a_, inc, hw in planes:
        nn = 24
        flags = np.zeros((nn, nn))
        for row in range(nn):
            be = hw - (row + 0.5) * 2 * hw / nn
            for col in range(nn):
                al = -hw + (col + 0.5) * 2 * hw / nn
                out = np.zeros(2)
                L.ref_eht_trace(a_, inc, al, be, ptr(out))
                flags[row, col] = out[0]
        cap.append(flags)
    g["eht_planes"] = np.array(planes)
    g["eht_captured"] = np.array(cap)

    # critical curves (eht.cpp:52-97)
    curves = [(0.9, math.radians(17.0)), (0.9, math.radians(60.0)), (0.0, math.radians(40.0)),
              (0.998, math.radians(85.0)), (0.5, math.radians(30.0))]
    pts, radii = [], []
    for a_, inc in curves:
        al = np.zeros(4000)
        be = np.zeros(4000)
        k = L.ref_critical_curve(a_, inc, 100, ptr(al), ptr(be), 4000)
        pts.append(np.stack([al[:k], be[:k]], axis=1))
        psis = 2 * math.pi * (np.arange(16) + 0.25) / 16
        radii.append([L.ref_critical_radius(ptr(al), ptr(be), k, p) for p in psis])
    g["curve_params"] = np.array(curves)
    for j, p in enumerate(pts):
        g[f"curve_{j}"] = p
    g["curve_radii"] = np.array(radii)

    # trajectory::traceRay known answers of the reference's testRayTrace (test_kerr.cpp:496-543)
    cases = []
    r_far = 1000.0
    bc = 3.0 * math.sqrt(3.0)

    def look_for(b):
        sa = b * math.sqrt(1.0 - 2.0 / r_far) / r_far
        return [-math.sqrt(1.0 - sa * sa), 0.0, sa]

    cases.append((0.9, 30.0, math.pi / 2, [-1.0, 0.0, 0.0], 0, 2.0, 20.0, 250.0))     # at the hole
    cases.append((0.9, 30.0, math.pi / 2, [0.0, 0.0, 1.0], 0, 2.0, 20.0, 250.0))      # sideways
    cases.append((0.0, r_far, math.pi / 2, look_for(bc * (1 + 1e-6)), 0, 2.0, 20.0, 250.0))  # grazing
    cases.append((0.0, r_far, math.pi / 2, look_for(2 * bc), 0, 2.0, 20.0, 250.0))          # wide
    risco = L.ref_isco_radius(0.9, 1)
    cases.append((0.9, 30.0, 1.35, [-0.95, 0.3122, 0.0], 1, risco, 20.0, 250.0))     # disk
    tr = np.zeros((len(cases), 11))
    for k, (a_, r_, th_, lk, disk, rin, rout, esc) in enumerate(cases):
        lkv = np.ascontiguousarray(np.array(lk, dtype=float))
        out = np.zeros(11)
        L.ref_trajectory_trace(a_, r_, th_, 0.0, ptr(lkv), disk, rin, rout, esc, ptr(out))
        tr[k] = out
    g["traj_cases"] = np.array([[c[0], c[1], c[2], *c[3], c[4], c[5], c[6], c[7]] for c in cases])
    g["traj_out"] = tr

    # exact Schwarzschild bending integral (central.cpp:388-408)
    ba = np.array([[R, math.radians(d)] for R in (5.8, 8.0, 20.0) for d in (10.0, 40.0, 70.0, 89.0)])
    g["bend_in"] = ba
    g["bend"] = np.array([L.ref_bending_angle(R, al) for R, al in ba])

    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT}: {len(g)} arrays")


if __name__ == "__main__":
    sys.exit(main())
