"""MI355X-native Kerr null-geodesic tracer (HIP for gfx950 behind a C ABI).

The product is libkerr_trace.so, built in-tree from csrc/ (see build.py); this
package loads it with ctypes.  Public names mirror the reference's entry points.
"""
from .api import (REPORT_FIELDS, TraceOpts, Tracer, View, blackbody_table, circular_omega, critical_curve, critical_radius,
                  geodesic_four_velocity, horizon_radius, observer_frame, isco_radius, photon_orbit_radius, render, shadow_radius, to_numpy, trace_looks,
                  trace_ray)
from ._lib import (CAMERA_BARDEEN, CAMERA_OBSERVER, CAMERA_PINHOLE, CAPTURED, DISK, ESCAPED, FLAG_CLOSED_FORM, FLAG_FP64_RESTEP, FLAG_SURROGATE,
                   MODE_AUTO, MODE_ONESHOT, MODE_PERSISTENT, NET_FAST, NET_WIDE, PARTICLE_ESCAPED, PARTICLE_MOVING, PARTICLE_PLUNGED,
                   PARTICLE_UNFINISHED, PREC_FP32, PREC_FP64, PREC_MIXED, SKY_BLACK, SKY_GRID,
                   SKY_IMAGE, UNFINISHED)

__all__ = [
    "REPORT_FIELDS", "TraceOpts", "Tracer", "View", "blackbody_table", "circular_omega", "critical_curve", "critical_radius",
    "geodesic_four_velocity", "horizon_radius", "observer_frame", "isco_radius", "photon_orbit_radius", "render", "shadow_radius", "to_numpy", "trace_looks",
    "trace_ray",
    "CAMERA_BARDEEN", "CAMERA_OBSERVER", "CAMERA_PINHOLE", "CAPTURED", "DISK", "ESCAPED", "FLAG_CLOSED_FORM", "FLAG_FP64_RESTEP", "FLAG_SURROGATE",
    "MODE_AUTO", "MODE_ONESHOT", "MODE_PERSISTENT", "NET_FAST", "NET_WIDE", "PARTICLE_ESCAPED", "PARTICLE_MOVING", "PARTICLE_PLUNGED",
    "PARTICLE_UNFINISHED", "PREC_FP32", "PREC_FP64", "PREC_MIXED", "SKY_BLACK", "SKY_GRID",
    "SKY_IMAGE", "UNFINISHED",
]
