"""The C++20 host face (include/kerr_trace.hpp) through the kt_inspect program: the same
rays as the oracle, bit for bit, as the Python path."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

OL = O.lib()
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "black_simulation_amd", "kt_inspect")


def run(*args):
    out = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_inspect_single_pixels_match_oracle():
    """One pixel's full report, as App::inspectPixel prints it (src/app.cpp:776-777): the
    end, r_end, g, turns, r_min, delay, b and q."""
    v = O.make_view(width=64, height=64)
    for col, row in [(32, 32), (0, 0), (50, 32), (20, 33), (32, 20)]:
        got = run(0.9, 30.0, 80.0, 12.0, 64, 64, col, row)
        want = O.trace_pixel(v, O.make_opts(), col, row)
        assert got["end"] == want.cls and got["steps"] == want.steps
        rep = O.trace_params(v, O.make_opts(), [[got["alpha"], got["beta"]]], report=True)
        for k in ("turns", "r_min", "delay"):
            assert got[k] == pytest.approx(rep[k][0], rel=1e-12, abs=1e-13), k
        assert got["b"] == rep["b"][0] and got["q"] == rep["q"][0]
        r_end = {O.ESCAPED: 1000.0, O.CAPTURED: 1.01 * OL.ko_horizon_radius(0.9), O.DISK: want.r_hit}[want.cls]
        assert got["r_end"] == pytest.approx(r_end, rel=1e-15)
        if want.cls == O.ESCAPED:
            assert abs(got["theta_inf"] - want.theta_inf) < 1e-13
        else:
            assert got["theta_inf"] is None
        if want.cls == O.DISK:
            assert got["r_hit"] == want.r_hit and got["g"] == want.g


def test_render_matches_oracle_counts():
    got = run(0.9, 30.0, 80.0, 12.0, 48, 48)
    want = O.render(O.make_view(width=48, height=48), O.make_opts())
    counts = np.bincount(want["cls"], minlength=4)
    assert [got["escaped"], got["captured"], got["disk"], got["unfinished"]] == counts.tolist()
    assert got["steps"] == int(want["steps"].sum())


def test_eht_render_and_cancel_through_the_cpp_face():
    """kt::Tracer::ehtRender (eht::render's shape): the image's total equals the Python
    path's, and a cancel flag another thread sets before the render begins (the tracer's
    creation takes longer than the thread's wait) leaves it untraced and cancelled."""
    from black_simulation_amd import eht

    got = run("eht", 0.9, 17.0, 16.0, 64)
    assert got["cancelled"] is False and got["rows_with_light"] > 0
    img = eht.render(eht.EhtView(0.9, math.radians(17.0), 16.0, 64))
    want = float(img.intensity.sum().item())
    assert got["intensity_sum"] == pytest.approx(want, rel=1e-12)
    got = run("eht", 0.9, 17.0, 16.0, 64, 0)
    assert got["cancelled"] is True and got["rows_with_light"] == 0 and got["intensity_sum"] == 0.0
