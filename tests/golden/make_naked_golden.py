"""Writes tests/golden/naked_vectors.npz: the reference's own outputs past a = 1 (a naked
singularity, src/kerr.cpp:117 and its naked branches), from the reference's code compiled
in place into oracle/_ref/libkerr_ref.so (oracle/Makefile `ref`), as make_golden.py does.

- the radii (kerr::horizonRadius, iscoRadius, photonOrbitRadius) at five spins in (1, 1.5];
- kerr::azimuthShift on both sides of r = 0;
- eht::traceRay's capture flags and least radii over three Bardeen planes from r = 1000 M
  (a ray is captured where it passes through the ring's disk to r < 0: src/eht.cpp:285);
- trajectory::traceRay outcomes of rays from a ZAMO camera, with the disk.

The fixture holds only inputs and the reference's outputs.

    python tests/golden/make_naked_golden.py
"""
from __future__ import annotations

import math
import os

import numpy as np

from make_golden import load_ref, ptr

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "naked_vectors.npz")


def main():
    L = load_ref()
    g = {}
    spins = np.array([1.0001, 1.05, 1.2, 1.35, 1.5])
    g["spins"] = spins
    g["radii"] = np.array([[L.ref_horizon_radius(a), L.ref_isco_radius(a, 1), L.ref_isco_radius(a, 0),
                            L.ref_photon_orbit_radius(a, 1), L.ref_photon_orbit_radius(a, 0)] for a in spins])
    g["azshift_in"] = np.array([[a, r] for a in (1.05, 1.2, 1.5) for r in (-6.0, -1.0, -0.2, 0.0, 0.3, 1.0, 2.5, 9.0,
                                                                          40.0)])
    g["azshift"] = np.array([L.ref_azimuth_shift(a, r) for a, r in g["azshift_in"]])

    planes = [(1.2, math.radians(60.0), 2.5), (1.5, math.radians(17.0), 2.5), (1.05, math.radians(85.0), 2.0)]
    cap, rmin = [], []
    for a_, inc, hw in planes:
        nn = 24
        flags, least = np.zeros((nn, nn)), np.zeros((nn, nn))
        for row in range(nn):
            be = hw - (row + 0.5) * 2 * hw / nn
            for col in range(nn):
                al = -hw + (col + 0.5) * 2 * hw / nn
                out = np.zeros(2)
                L.ref_eht_trace(a_, inc, al, be, ptr(out))
                flags[row, col], least[row, col] = out
        cap.append(flags)
        rmin.append(least)
    g["eht_planes"] = np.array(planes)
    g["eht_captured"] = np.array(cap)
    g["eht_rmin"] = np.array(rmin)

    # a fan of looks from a ZAMO camera at r = 30 M, i = 77 deg, around the direction of the
    # hole (look = -e_r), with the prograde disk from the reference's ISCO to 20 M
    a_, r_, th_ = 1.2, 30.0, 1.35
    risco = L.ref_isco_radius(a_, 1)
    cases = []
    for dy in np.linspace(-0.12, 0.12, 7):
        for dz in np.linspace(-0.12, 0.12, 7):
            lk = np.array([-1.0, dy, dz])
            cases.append(lk / np.linalg.norm(lk))
    tr = np.zeros((len(cases), 11))
    for k, lk in enumerate(cases):
        lkv = np.ascontiguousarray(lk)
        out = np.zeros(11)
        L.ref_trajectory_trace(a_, r_, th_, 0.0, ptr(lkv), 1, risco, 20.0, 250.0, ptr(out))
        tr[k] = out
    g["traj_camera"] = np.array([a_, r_, th_, risco, 20.0, 250.0])
    g["traj_looks"] = np.array(cases)
    g["traj_out"] = tr

    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT}: {len(g)} arrays")


if __name__ == "__main__":
    main()
