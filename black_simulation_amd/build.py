"""Builds libkerr_trace.so in-tree with hipcc for gfx950 (no JIT, no site-packages)."""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["kt_host.cpp", "kt_trace.hip", "kt_nn.hip", "kt_surrogate.hip", "kt_shade.hip", "kt_eht.hip", "kt_particles.hip"]

# -ffp-contract=off: the FP64 step loop must round exactly as written (its results are
# bit-identical to any IEEE implementation doing the same operations).  Packed FP32 is
# disabled on purpose.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
]


def build(verbose: bool = False) -> str:
    out = os.path.join(PKG_DIR, "libkerr_trace.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = [ln for ln in (res.stderr or "").splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if res.returncode != 0:
        sys.stderr.write("\n".join(noise) + "\n")
        raise RuntimeError("hipcc failed building libkerr_trace.so")
    os.replace(tmp, out)
    # the C++20 inspector, linked against the library through include/kerr_trace.hpp
    exe = os.path.join(PKG_DIR, "kt_inspect")
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-o", exe + ".tmp", os.path.join(CSRC, "kt_inspect.cpp"),
           "-L" + PKG_DIR, "-lkerr_trace", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stderr)
        raise RuntimeError("g++ failed building kt_inspect")
    os.replace(exe + ".tmp", exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
