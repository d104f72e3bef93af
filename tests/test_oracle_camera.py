"""The oracle's pinhole camera (camera.cpp:6-21, trace.frag:1495-1496) against what the
geometry of the camera fixes, not against a restatement of its formula.

Every check launches a pixel through the oracle (ko_launch_pixel) and reads the look
direction back from the launched photon alone, in the static observer's orthonormal
frame at a = 0 (where the ZAMO is the static observer):

    f = 1 - 2/r,  E_loc = E/sqrt(f),  p_(r) = sqrt(f) p_r,  p_(theta) = p_theta/r,
    p_(phi) = L/(r sin theta),  p_r = -w_r/Delta,  p_theta = w_mu/sin theta,

with (w_r, w_mu) the backward Mino-time rates of the launch state.  The camera looks
along n = -p_(i)/E_loc: against the arriving photon's motion.  The expected directions
are the camera's stated axes (forward = -e_r, up = -e_theta = north, right = +e_phi),
what yaw, pitch and roll mean (turn toward the right, tilt toward up, turn the image
about the view axis), pixel centres on a square grid scaled by tan(fov/2) over half the
height, and row 0 at the top.  A flipped yaw, pitch or roll sign, swapped right and up,
a transposed basis, a bottom-up row order or a wrong field of view fails one of them."""
import math

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
R_CAM = 30.0
DEG = math.pi / 180.0


def launch(view, col, row):
    p = O.Photon()
    y0 = (O.C.c_double * 5)()
    L.ko_launch_pixel(O.C.byref(view), col, row, O.C.byref(p), y0)
    return p, np.array(y0)


def look(col, row, *, width=21, height=21, theta=math.pi / 2, **cam):
    """Unit look direction (r, theta, phi components) of a pixel of a static camera at
    a = 0, recovered from the photon the oracle launches for it."""
    v = O.make_view(camera_kind=O.PINHOLE, width=width, height=height, spin=0.0, r_cam=R_CAM,
                    theta_cam=theta, **cam)
    p, y = launch(v, col, row)
    r, mu, _, w_r, w_mu = y
    sn = math.sqrt(1.0 - mu * mu)
    f = 1.0 - 2.0 / r
    delta = r * r - 2.0 * r
    e_loc = p.E / math.sqrt(f)
    p_r = -w_r / delta
    p_th = w_mu / sn
    d = np.array([math.sqrt(f) * p_r, p_th / r, p.L / (r * sn)])
    return -d / e_loc


def test_look_direction_is_a_unit_vector():
    """The read-back is itself checked: a null photon gives |n| = 1 for any pixel."""
    rng = np.random.default_rng(3)
    for _ in range(30):
        col, row = rng.integers(0, 21, size=2)
        n = look(int(col), int(row), yaw=rng.uniform(-3, 3), pitch=rng.uniform(-1.5, 1.5),
                 roll=rng.uniform(-3, 3), theta=rng.uniform(0.3, 2.8))
        assert np.linalg.norm(n) == pytest.approx(1.0, abs=1e-13)


@pytest.mark.parametrize("yaw,pitch,want", [
    (0.0, 0.0, (-1, 0, 0)),            # forward: at the hole
    (90.0, 0.0, (0, 0, 1)),            # yaw toward the right (+e_phi)
    (-90.0, 0.0, (0, 0, -1)),
    (180.0, 0.0, (1, 0, 0)),           # away from the hole
    (0.0, 90.0, (0, -1, 0)),           # pitch toward up: north, -e_theta
    (0.0, -90.0, (0, 1, 0)),
    (90.0, 30.0, (0, -0.5, math.sqrt(3) / 2)),  # yawed right, then tilted up
    (-45.0, -30.0, (-math.sqrt(3) / 2 / math.sqrt(2), 0.5, -math.sqrt(3) / 2 / math.sqrt(2))),
])
def test_centre_pixel_looks_along_yawed_and_pitched_forward(yaw, pitch, want):
    n = look(10, 10, yaw=yaw * DEG, pitch=pitch * DEG)
    np.testing.assert_allclose(n, want, atol=1e-13)


@pytest.mark.parametrize("roll", [0.0, 37.0, 90.0, -120.0])
def test_roll_leaves_the_centre_pixel_alone(roll):
    np.testing.assert_allclose(look(10, 10, yaw=0.4, pitch=0.3, roll=roll * DEG),
                               look(10, 10, yaw=0.4, pitch=0.3), atol=1e-13)


def test_row_zero_is_the_top_and_column_zero_the_left():
    """Facing the hole at the equator: the top row looks north (-e_theta), the bottom row
    south, the first column toward -e_phi (left), the last toward +e_phi (right)."""
    top, bottom = look(10, 0), look(10, 20)
    left, right = look(0, 10), look(20, 10)
    assert top[1] < -0.1 and bottom[1] > 0.1
    assert abs(top[2]) < 1e-13 and abs(bottom[2]) < 1e-13
    assert left[2] < -0.1 and right[2] > 0.1
    assert abs(left[1]) < 1e-13 and abs(right[1]) < 1e-13


@pytest.mark.parametrize("width,height,fov", [(21, 21, 60.0), (33, 15, 90.0), (8, 6, 40.0)])
def test_field_of_view_and_square_pixels(width, height, fov):
    """A pixel centre at (x, y) pixels from the image centre is off the view axis by
    atan(tan(fov/2) * hypot(x, y) / (height/2)): square pixels, fov over the height."""
    cx, cy = 0.5 * width, 0.5 * height
    axis = np.array([-1.0, 0.0, 0.0])
    t = math.tan(0.5 * fov * DEG)
    for col, row in [(0, 0), (width - 1, 0), (0, height - 1), (width - 1, height - 1),
                     (width // 3, height // 2), (width // 2, 1)]:
        n = look(col, row, width=width, height=height, fov_deg=fov)
        off = math.hypot(col + 0.5 - cx, row + 0.5 - cy)
        assert math.acos(np.clip(n @ axis, -1, 1)) == pytest.approx(math.atan(t * off / cy), abs=1e-12)
        # and on the side the pixel sits: right of centre looks toward +e_phi, above toward -e_theta
        if col + 0.5 != cx:
            assert np.sign(n[2]) == np.sign(col + 0.5 - cx)
        if row + 0.5 != cy:
            assert np.sign(-n[1]) == np.sign(cy - row - 0.5)


@pytest.mark.parametrize("yaw,pitch", [(0.0, 0.0), (70.0, 20.0), (-130.0, -55.0), (10.0, 80.0)])
def test_no_roll_keeps_the_horizon_level(yaw, pitch):
    """Without roll the image's horizontal axis has no north-south part: pixels mirrored
    about the centre within the middle row look equally far north."""
    for k in (1, 4, 10):
        a = look(10 - k, 10, yaw=yaw * DEG, pitch=pitch * DEG)
        b = look(10 + k, 10, yaw=yaw * DEG, pitch=pitch * DEG)
        assert a[1] == pytest.approx(b[1], abs=1e-13)
        assert a[1] != pytest.approx(0.0, abs=1e-3) or pitch == 0.0


@pytest.mark.parametrize("roll,map_to", [
    (90.0, lambda k: (10, 10 - k)),    # the pixel k right of centre sees what k above did
    (-90.0, lambda k: (10, 10 + k)),   # ... or k below
    (180.0, lambda k: (10 - k, 10)),   # ... or k left
])
def test_roll_turns_right_toward_up(roll, map_to):
    """Roll turns the camera's right axis toward its up axis about the view direction, so
    the image content turns the other way: with roll +90 the pixel right of centre sees
    what the unrolled camera saw above centre."""
    for k in (1, 3, 10):
        col, row = map_to(k)
        np.testing.assert_allclose(look(10 + k, 10, yaw=0.5, pitch=-0.2, roll=roll * DEG),
                                   look(col, row, yaw=0.5, pitch=-0.2), atol=1e-13)


def test_roll_by_an_arbitrary_angle_rotates_pixel_offsets():
    """A pixel at offset (x, y) from the centre (y up), under roll psi, looks where the
    unrolled camera's pixel at (x cos psi - y sin psi, x sin psi + y cos psi) does.  The
    rotated offset is not a pixel centre, so it is compared through the camera's own
    direction rule with the unrolled camera's axes read off three pixels."""
    psi = 27.0 * DEG
    t = math.tan(30.0 * DEG)
    f = look(10, 10, yaw=0.3, pitch=0.2)
    right = look(11, 10, yaw=0.3, pitch=0.2) / (look(11, 10, yaw=0.3, pitch=0.2) @ f) - f
    up = look(10, 9, yaw=0.3, pitch=0.2) / (look(10, 9, yaw=0.3, pitch=0.2) @ f) - f
    for col, row in [(14, 10), (10, 3), (2, 17), (19, 0)]:
        x, y = (col + 0.5 - 10.5) / 10.5 * t, (10.5 - row - 0.5) / 10.5 * t
        xs, ys = x * math.cos(psi) - y * math.sin(psi), x * math.sin(psi) + y * math.cos(psi)
        s = 10.5 / t  # the right/up vectors read off adjacent pixels are per pixel
        d = f + (xs * s) * right + (ys * s) * up
        np.testing.assert_allclose(look(col, row, yaw=0.3, pitch=0.2, roll=psi), d / np.linalg.norm(d), atol=1e-12)


def test_centre_ray_of_a_resting_camera_is_radial_and_captured():
    """At a = 0 the centre pixel of an odd-sized view at zero yaw, pitch and roll carries
    L = 0 and Q = 0 (a radial photon) and falls into the hole; at yaw 180 it is radial
    and escapes straight out, to the camera's own sky position."""
    v = O.make_view(camera_kind=O.PINHOLE, width=21, height=21, spin=0.0, r_cam=R_CAM, theta_cam=1.1, phi_cam=0.7)
    p, _ = launch(v, 10, 10)
    assert abs(p.L) < 1e-14 and abs(p.Q) < 1e-12
    res = O.trace_pixel(v, O.make_opts(), 10, 10)
    assert res.cls == O.CAPTURED
    v.yaw = math.pi
    res = O.trace_pixel(v, O.make_opts(), 10, 10)
    assert res.cls == O.ESCAPED
    assert res.theta_inf == pytest.approx(1.1, abs=1e-9)
    assert res.phi_inf == pytest.approx(0.7, abs=1e-9)


def test_zamo_centre_ray_has_zero_angular_momentum_at_any_spin():
    """The ZAMO's axes have no e_phi part along e_r, so a photon along -e_r of the ZAMO
    frame has L = p_phi = 0 for every spin."""
    for a in (0.5, 0.9, 0.998):
        v = O.make_view(camera_kind=O.PINHOLE, width=21, height=21, spin=a, r_cam=R_CAM, theta_cam=1.2)
        p, _ = launch(v, 10, 10)
        assert abs(p.L) < 1e-13


def test_yaw_sign_from_where_rays_leave():
    """Traced back, the centre ray of a camera yawed right (+90) leaves toward +phi and
    escapes at a sky longitude in (0, pi) from the camera; yawed left it escapes in
    (pi, 2 pi).  Pitched up (+60) at the equator it escapes into the northern sky."""
    v = O.make_view(camera_kind=O.PINHOLE, width=21, height=21, spin=0.0, r_cam=R_CAM, theta_cam=math.pi / 2)
    v.yaw = math.pi / 2
    res = O.trace_pixel(v, O.make_opts(), 10, 10)
    assert res.cls == O.ESCAPED and 0.0 < res.phi_inf < math.pi
    v.yaw = -math.pi / 2
    res = O.trace_pixel(v, O.make_opts(), 10, 10)
    assert res.cls == O.ESCAPED and math.pi < res.phi_inf < 2 * math.pi
    v.yaw, v.pitch = 0.0, 60.0 * DEG
    res = O.trace_pixel(v, O.make_opts(), 10, 10)
    assert res.cls == O.ESCAPED and res.theta_inf < math.pi / 2
