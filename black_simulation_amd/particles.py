"""Massive particles (debris) on Kerr geodesics, on the GPU: the geodesic part of the
simulator's debris update (shaders/particles_update.comp:207-255) and of
kerr::stepCoordinateTime (src/kerr.cpp:544-561) over libkerr_trace.so.

    p = Particles.from_zamo(a, r, theta, phi, v)     ~ kerr::launchFromZamo, mu = 1
    p.advance(tracer, a, dt)                          ~ one debris update of dt (Kerr-Schild time)

The state stays on the device in the integrator's variables: r, mu = cos(theta), the
Kerr-Schild phi, vr = Sigma dr/dtau and vmu = Sigma dmu/dtau (= -sin(theta) vtheta), with
the constants E, L, Q of each particle and its fate.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as K
from .api import TraceOpts, Tracer, _check

COLUMNS = ("r", "mu", "phi", "vr", "vmu", "E", "L", "Q")


class Particles:
    """n particles as device tensors (float64 columns, int8 fate, int32 steps)."""

    def __init__(self, columns: dict, device: int = 0):
        import torch

        self.torch = torch
        dev = f"cuda:{device}"
        n = len(columns["r"])
        self.n = n
        for c in COLUMNS:
            setattr(self, c, torch.as_tensor(np.ascontiguousarray(columns[c], np.float64)).to(dev))
        fate = columns.get("fate")
        self.fate = (torch.zeros(n, dtype=torch.int8) if fate is None else torch.as_tensor(
            np.ascontiguousarray(fate, np.int8))).to(dev)
        self.steps = torch.zeros(n, dtype=torch.int32, device=dev)
        self.h = torch.zeros(n, dtype=torch.float64, device=dev)  # the controller's step, carried over

    @classmethod
    def from_zamo(cls, a: float, r, theta, phi, v, device: int = 0) -> "Particles":
        """Particles at (r, theta, phi) with 3-velocities v (n x 3, along e_r, e_theta, e_phi)
        relative to the local ZAMO (kt_particle_from_zamo, kerr::launchFromZamo with mu = 1)."""
        r, theta, phi = (np.atleast_1d(np.asarray(x, np.float64)) for x in (r, theta, phi))
        v = np.atleast_2d(np.asarray(v, np.float64))
        n = len(r)
        cols = {c: np.zeros(n) for c in COLUMNS}
        out = (C.c_double * 5)()
        lib = K.lib()
        for i in range(n):
            vi = (C.c_double * 3)(*v[i])
            if lib.kt_particle_from_zamo(a, float(r[i]), float(theta[i]), vi, out) != K.KT_OK:
                raise ValueError(f"particle {i}: no ZAMO at r = {r[i]} (inside the horizon?)")
            cols["E"][i], cols["L"][i], cols["Q"][i], cols["vr"][i], cols["vmu"][i] = out
        cols["r"], cols["mu"], cols["phi"] = r, np.cos(theta), phi
        return cls(cols, device)

    def _struct(self) -> K.KtParticles:
        return K.KtParticles(*[getattr(self, c).data_ptr() for c in COLUMNS], self.fate.data_ptr(),
                             self.steps.data_ptr(), self.h.data_ptr(), 1)

    def advance(self, tracer: Tracer, a: float, dt: float, r_stop: float = 0.0, r_escape: float = 1e4,
                opts: TraceOpts | None = None, stream=None) -> "Particles":
        """kt_particles_advance: every moving particle by dt of Kerr-Schild coordinate time
        (asynchronous on `stream`)."""
        st = stream if stream is not None else self.torch.cuda.current_stream(tracer.device)
        s = self._struct()
        o = (opts or TraceOpts()).to_c()
        _check(tracer._ctx, K.lib().kt_particles_advance(tracer._ctx, a, C.byref(s), self.n, dt, r_stop, r_escape,
                                                         C.byref(o), C.c_void_p(st.cuda_stream)),
               "kt_particles_advance")
        return self

    def to_numpy(self) -> dict:
        self.torch.cuda.synchronize()
        out = {c: getattr(self, c).cpu().numpy() for c in COLUMNS}
        out["fate"] = self.fate.cpu().numpy()
        out["steps"] = self.steps.cpu().numpy()
        out["h"] = self.h.cpu().numpy()
        return out
