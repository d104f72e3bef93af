"""The learned surrogate's GPU kernels against the FP64 numpy network (oracle/mlp_oracle.py):
forward outputs, loss, gradients and one Adam step; then a short training run on
integrator-made data and the gated trace with its integrator fallback."""
import math

import numpy as np
import pytest
import torch

import black_simulation_amd as bsa
from black_simulation_amd.configs import c2_view, c5_view
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def bf16(x):
    """round to bfloat16 (nearest even), as the kernels store weights and activations"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def flat(params):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in params]).astype(np.float32)


def unflat(p, sizes=M.SIZES):
    out, o = [], 0
    for n_in, n_out in zip(sizes[:-1], sizes[1:]):
        W = p[o:o + n_in * n_out].reshape(n_in, n_out).astype(np.float64)
        o += n_in * n_out
        b = p[o:o + n_out].astype(np.float64)
        o += n_out
        out.append((W, b))
    return out


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def quantised(params):
    """the weights as the kernels use them: FP32 input layer, bf16 matrix-core layers"""
    return [(f32(W) if l == 0 else bf16(W), b) for l, (W, b) in enumerate(params)]


def emulated_forward(params, x):
    """the kernels' arithmetic: FP32 input layer, then bf16 weights and layer inputs with
    wide accumulation"""
    Wq = quantised(params)
    h = f32(x)
    for l, (W, b) in enumerate(Wq):
        z = h @ W + b
        h = bf16(M.silu(z)) if l < len(Wq) - 1 else z
    return h, Wq


NEG_LOG2E = np.float32(-1.44269504088896341)
NEG_LN2 = np.float32(-0.693147180559945309)


def fused_forward(params, x):
    """the fused inference kernel's arithmetic: pre-activations scaled by -log2(e) (so
    exp(-z) = 2^z'), activations stored as bf16(-log2(e) silu(z)), the input layer as
    split-bf16 products fh Wh + fl Wh + fh Wl, bf16 weights, the output times -ln(2)"""
    (W0, b0), hidden, (W4, b4) = params[0], params[1:-1], params[-1]
    scaled = lambda v: (np.asarray(v, np.float32) * NEG_LOG2E).astype(np.float64)  # float32 products
    W0s = scaled(W0)
    Wh = bf16(W0s)
    Wl = bf16(W0s - Wh)
    xf = f32(x)
    fh = bf16(xf)
    fl = bf16(xf - fh)
    z = fh @ Wh + fl @ Wh + fh @ Wl + scaled(b0)
    act = lambda zs: bf16(zs / (1.0 + np.exp2(zs)))
    h = act(z)
    for W, b in hidden:
        h = act(h @ bf16(W) + scaled(b))
    return float(NEG_LN2) * (h @ bf16(W4)) + f32(b4)


def problem(seed=3, n=1024, sizes=M.SIZES):
    rng = np.random.default_rng(seed)
    params = M.init_params(seed, sizes)
    params = [(W, rng.normal(0, 0.05, b.shape)) for W, b in params]
    x = rng.uniform(-1, 1, (n, 5))
    cls = rng.integers(0, 3, n).astype(np.int8)
    tgt = rng.uniform(-1, 1, (n, 5))
    return params, x, cls, tgt


def test_forward_matches_numpy(tracer):
    """Against the network (exact SiLU, FP32 input layer, bf16 weights and activations) to
    the bf16 noise, and against the fused kernel's own arithmetic far tighter: what is left
    there is FP32 accumulation order and the hardware exp2/rcp, which move an activation's
    bf16 rounding now and then."""
    params, x, _, _ = problem()
    tracer.surrogate_load(flat(params))
    y = tracer.surrogate_predict(x).cpu().numpy()
    P = unflat(flat(params))
    want, _ = emulated_forward(P, x)
    scale = np.abs(want).max()
    err = np.abs(y - want)
    print(f"forward: max err {err.max() / scale:.2e}, median {np.median(err) / scale:.2e} (relative to {scale:.2f})")
    assert err.max() < 2e-2 * scale and np.median(err) < 1e-3 * scale
    exact = fused_forward(P, x)
    err = np.abs(y - exact)
    print(f"fused arithmetic: max err {err.max() / scale:.2e}, median {np.median(err) / scale:.2e}")
    assert err.max() < 2e-3 * scale and np.median(err) < 2e-5 * scale


def test_gradients_match_numpy(tracer):
    params, x, cls, tgt = problem(n=2048)
    p32 = flat(params)
    tracer.surrogate_load(p32)
    g, loss = tracer.surrogate_gradients(x, cls, tgt)
    g = g.cpu().numpy()
    P = unflat(p32)
    Wq = quantised(P)
    # FP64 backward through the same network, activations rounded as on the GPU
    h = f32(x)
    cache = []
    for l, (W, b) in enumerate(Wq):
        z = h @ W + b
        cache.append((h, z))
        h = bf16(M.silu(z)) if l < len(Wq) - 1 else z
    want_loss, gy = M.loss_and_grad_out(h, cls.astype(int), tgt)
    grads = M.backward(Wq, cache, gy)
    assert loss == pytest.approx(want_loss, rel=2e-3)
    o = 0
    for l, (gW, gb) in enumerate(grads):
        for ref in (gW.ravel(), gb):
            got = g[o:o + ref.size]
            o += ref.size
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            cos = np.dot(got, ref) / (np.linalg.norm(got) * np.linalg.norm(ref))
            print(f"layer {l} {'W' if ref.ndim else 'b'} size {ref.size}: rel err {rel:.2e} cos {cos:.6f}")
            assert rel < 3e-2 and cos > 0.999


def test_adam_step_matches_numpy(tracer):
    params, _, _, _ = problem()
    p32 = flat(params)
    tracer.surrogate_load(p32)
    rng = np.random.default_rng(9)
    gr = rng.normal(0, 1e-3, p32.size).astype(np.float32)
    tracer.surrogate_adam_step(torch.as_tensor(gr).cuda(), 1e-3, 1)
    got = tracer.surrogate_save()
    st = M.adam_state(unflat(p32))
    want = flat(M.adam_step(unflat(p32), [(gW, gb) for gW, gb in unflat(gr)], st, 1e-3))
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7)


def f16(a):
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def fast_forward(params, x):
    """the fast kernel's arithmetic: activations kept divided by 6 and computed on packed
    halves (x = f16(z'), f16(x f16clamp(x + 1/2)) with z' = z / 6 in FP32), the input layer
    as split-bf16 products of W0 / 6, FP16 hidden and output weights, the output times 6
    plus b4"""
    (W0, b0), hidden, (W4, b4) = params[0], params[1:-1], params[-1]
    sixth = lambda v: (np.asarray(v, np.float32) * np.float32(1.0 / 6.0)).astype(np.float64)  # noqa: E731
    W0s = sixth(W0)
    Wh = bf16(W0s)
    Wl = bf16(W0s - Wh)
    xf = f32(x)
    fh = bf16(xf)
    fl = bf16(xf - fh)

    def act(zs):
        xh = f16(f32(zs))
        return f16(xh * np.clip(f16(xh + 0.5), 0.0, 1.0))

    h = act(fh @ Wh + fl @ Wh + fh @ Wl + sixth(b0))
    for W, b in hidden:
        h = act(h @ f16(W) + sixth(b))
    return 6.0 * (h @ f16(W4)) + f32(b4)


def test_fast_network_forward_matches_numpy(tracer):
    """The 128-wide hard-SiLU network in the fast kernel: against the network itself (FP32
    input layer, bf16 weights and activations, as trained) to the bf16 noise, and against the
    kernel's own arithmetic (activations divided by 6 on packed halves, FP16 hidden and
    output weights, split-bf16 input layer) far tighter."""
    params, x, _, _ = problem(n=2048, sizes=M.FAST_SIZES)
    tracer.surrogate_load(flat(params))
    assert tracer.surrogate_net() == bsa.NET_FAST
    y = tracer.surrogate_predict(x).cpu().numpy()
    P = unflat(flat(params), M.FAST_SIZES)
    Wq = quantised(P)
    h = f32(x)
    for l, (W, b) in enumerate(Wq):
        z = h @ W + b
        h = bf16(M.hsilu(z)) if l < len(Wq) - 1 else z
    scale = np.abs(h).max()
    err = np.abs(y - h)
    print(f"fast forward: max err {err.max() / scale:.2e}, median {np.median(err) / scale:.2e} (of {scale:.2f})")
    assert err.max() < 2e-2 * scale and np.median(err) < 1e-3 * scale
    exact = fast_forward(P, x)
    err = np.abs(y - exact)
    print(f"fast kernel arithmetic: max err {err.max() / scale:.2e}, median {np.median(err) / scale:.2e}")
    assert err.max() < 2e-3 * scale and np.median(err) < 2e-5 * scale


def test_fast_network_gradients_and_adam_match_numpy(tracer):
    """Training kernels on the fast network: loss and every layer's weight and bias
    gradient against the FP64 network (hard SiLU, activations rounded as on the GPU), then
    one Adam step."""
    params, x, cls, tgt = problem(n=2048, sizes=M.FAST_SIZES)
    p32 = flat(params)
    tracer.surrogate_load(p32)
    g, loss = tracer.surrogate_gradients(x, cls, tgt)
    g = g.cpu().numpy()
    P = unflat(p32, M.FAST_SIZES)
    Wq = quantised(P)
    h = f32(x)
    cache = []
    for l, (W, b) in enumerate(Wq):
        z = h @ W + b
        cache.append((h, z))
        h = bf16(M.hsilu(z)) if l < len(Wq) - 1 else z
    want_loss, gy = M.loss_and_grad_out(h, cls.astype(int), tgt)
    grads = M.backward(Wq, cache, gy, "hard")
    assert loss == pytest.approx(want_loss, rel=2e-3)
    o = 0
    for l, (gW, gb) in enumerate(grads):
        for ref in (gW.ravel(), gb):
            got = g[o:o + ref.size]
            o += ref.size
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            cos = np.dot(got, ref) / (np.linalg.norm(got) * np.linalg.norm(ref))
            print(f"fast layer {l} {'W' if ref.ndim else 'b'} size {ref.size}: rel err {rel:.2e} cos {cos:.6f}")
            assert rel < 3e-2 and cos > 0.999
    assert o == p32.size
    rng = np.random.default_rng(9)
    gr = rng.normal(0, 1e-3, p32.size).astype(np.float32)
    tracer.surrogate_adam_step(torch.as_tensor(gr).cuda(), 1e-3, 1)
    got = tracer.surrogate_save()
    st = M.adam_state(P)
    want = flat(M.adam_step(P, [(gW, gb) for gW, gb in unflat(gr, M.FAST_SIZES)], st, 1e-3))
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7)


@pytest.fixture(scope="module")
def trained(tracer):
    tracer.surrogate_init(7)
    rep = tracer.surrogate_train(n_train=1 << 17, n_holdout=1 << 14, epochs=6, batch=8192, seed=11)
    print("short training:", {k: round(v, 6) if isinstance(v, float) else v for k, v in rep.items()}, flush=True)
    return rep


def test_short_training_learns(trained):
    assert trained["final_loss"] < 0.7 * trained["first_loss"]
    assert trained["class_agreement"] > 0.85


def test_gated_trace_falls_back_to_the_integrator(tracer, trained):
    """Rays the gate rejects are traced by the FP64 integrator (bit-equal to the oracle);
    rays it accepts carry the surrogate flag and zero steps (an orbit frame: inside the
    network's domain)."""
    v = c5_view(60, 96, 54)
    # the briefly trained network is rarely 99.9 % sure: a looser gate lets it answer
    got = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(use_surrogate=1, gate_prob=0.9)))
    st = tracer.stats()
    ov = O.make_view(**{k: getattr(v, k) for k in (
        "camera_kind", "width", "height", "spin", "r_cam", "theta_cam", "phi_cam", "half_width", "half_height",
        "yaw", "pitch", "roll", "fov_deg", "disk_inner", "disk_outer", "disk_prograde", "disk_enabled",
        "escape_radius", "capture_factor")})
    want = O.render(ov, O.make_opts())
    sur = (got["cls"] & bsa.FLAG_SURROGATE) != 0
    fb = ~sur
    print(f"surrogate answered {sur.mean():.3f} of rays; fallback {st['fallback_rays']}", flush=True)
    assert sur.any() and fb.any()
    assert st["fallback_rays"] == fb.sum()
    assert np.array_equal(got["cls"][fb] & 7, want["cls"][fb])
    assert np.array_equal(got["steps"][fb], want["steps"][fb])
    assert np.all(got["steps"][sur] == 0)


def test_surrogate_plan_replays_equal_eager(tracer, trained):
    """Network, gate and fallback integrator captured in one hipGraph: replays with eager
    work and synchronisation in between reproduce the eager call every time."""
    v = c5_view(60, 96, 54)
    opts = bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED, gate_prob=0.9)
    eager = bsa.to_numpy(tracer.trace_views([v], opts))
    assert ((eager["cls"] & bsa.FLAG_SURROGATE) != 0).any()
    plan = tracer.plan([v], opts)
    for rep in range(3):
        out = plan.launch()
        torch.cuda.synchronize()
        got = bsa.to_numpy(out)
        for k in ("cls", "steps", "r_hit", "g", "theta_inf", "phi_inf"):
            assert np.array_equal(got[k], eager[k], equal_nan=True), (rep, k)
        tracer.trace_views([c2_view(32, 18)])  # unrelated eager work between replays
        torch.cuda.synchronize()
        for k in out:
            out[k].zero_()
    plan.close()


def test_trained_network_answers_most_rays_of_an_orbit_frame(tracer):
    """The shipped weights (black_simulation_amd/data/surrogate.f32) on a 1/16-size frame
    of the a=0.998 orbit: the gate answers most rays, classes agree with the FP64
    integrator, and the answered rays take no integrator steps."""
    from black_simulation_amd import orbit as OR
    from black_simulation_amd.configs import c5_view
    from black_simulation_amd.train_surrogate import load_params

    tracer.surrogate_load(load_params())
    v = c5_view(60, 480, 270)
    ref = tracer.trace_views([v], bsa.TraceOpts(precision=bsa.PREC_FP64))
    plain_steps = int(ref["steps"].to(torch.int64).sum())
    got = tracer.trace_views([v], bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED))
    e = OR.errors(got, ref)
    agree = float(e["agree"].float().mean())
    answered = float(e["answered"].float().mean())
    sur_steps = int(got["steps"].to(torch.int64).sum())
    d50, d99, _ = OR.quantiles(e["defl_answered"])
    print(f"answered {answered:.3f}, class agreement {agree:.5f}, integrator attempts {sur_steps} vs "
          f"{plain_steps} plain, deflection of answered escaped rays p50 {d50:.3g} p99 {d99:.3g} rad", flush=True)
    assert answered > 0.5
    assert agree > 0.99
    assert sur_steps < 0.7 * plain_steps


def test_surrogate_plans_replay_concurrently_on_two_streams(tracer, trained):
    """Two surrogate plans (each with its own outputs and fallback queue) replayed on two
    streams at once give what the eager calls give."""
    views = [c5_view(30, 96, 54), c5_view(100, 96, 54)]
    opts = bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED, gate_prob=0.9)
    eager = [bsa.to_numpy(tracer.trace_views([v], opts)) for v in views]
    assert all(((e["cls"] & bsa.FLAG_SURROGATE) != 0).any() for e in eager)
    streams = [torch.cuda.current_stream(), torch.cuda.Stream()]
    plans = [tracer.plan([v], opts) for v in views]
    for rep in range(3):
        outs = [p.launch(s) for p, s in zip(plans, streams)]
        torch.cuda.synchronize()
        for i, out in enumerate(outs):
            got = bsa.to_numpy(out)
            for k in ("cls", "steps", "r_hit", "g", "theta_inf", "phi_inf"):
                assert np.array_equal(got[k], eager[i][k], equal_nan=True), (rep, i, k)
            assert plans[i].stats()["fallback_rays"] == int(((eager[i]["cls"] & bsa.FLAG_SURROGATE) == 0).sum())
        for out in outs:
            for k in out:
                out[k].zero_()
        torch.cuda.synchronize()
    for p in plans:
        p.close()


def frame64(r, inc, phc, al, be):
    """the ray's straight-line frame (e1 undeflected direction, e2 towards the hole,
    e3 = e1 x e2), FP64, as the training targets use it"""
    si, ci, sp, cp = math.sin(inc), math.cos(inc), math.sin(phc), math.cos(phc)
    rh, th, ph = np.array([si * cp, si * sp, ci]), np.array([ci * cp, ci * sp, -si]), np.array([-sp, cp, 0.0])
    q = math.sqrt(max(1.0 - (al * al + be * be) / (r * r), 0.0))
    e1 = -q * rh - (be / r) * th - (al / r) * ph
    e1 /= np.linalg.norm(e1)
    e2 = -(r * rh - r * np.dot(rh, e1) * e1)
    e2 /= np.linalg.norm(e2)
    return e1, e2, np.cross(e1, e2)


def test_fused_features_and_sky_angles_match_fp64(tracer):
    """The fused kernel forms the network's inputs and the answered rays' sky angles in
    FP32.  Against the same steps in FP64 on the host (features through the shadow table,
    the network through surrogate_predict, the frame and the angles): the same classes,
    sky directions within 1e-5 rad at the median (a feature's last bit can move a bf16
    rounding inside the network, so a few rays differ more), r_hit and g likewise."""
    from black_simulation_amd.train_surrogate import load_params

    tracer.surrogate_load(load_params())
    # inside the square the network was trained on: +-1.5 mean shadow radii
    R = tracer.surrogate_shadow_radius(0.9, math.radians(85.0))
    v = bsa.View(width=96, height=96, spin=0.9, r_cam=40.0, theta_cam=math.radians(85.0), phi_cam=0.7,
                 half_width=1.45 * R, half_height=1.45 * R, disk_inner=0.0, disk_outer=20.0, escape_radius=1000.0,
                 capture_factor=1.01)
    got = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED)))
    sur = (got["cls"] & bsa.FLAG_SURROGATE) != 0
    assert sur.mean() > 0.5
    px, py = 2 * v.half_width / v.width, 2 * v.half_height / v.height
    idx = np.nonzero(sur)[0]
    al = -v.half_width + (idx % v.width + 0.5) * px
    be = v.half_height - (idx // v.width + 0.5) * py
    a, inc = v.spin, v.theta_cam
    R = tracer.surrogate_shadow_radius(a, inc)
    feat = np.stack([np.full(len(idx), math.log10(v.r_cam) - 2.0), np.full(len(idx), (math.cos(inc) - 0.5415) / 0.4545),
                     np.full(len(idx), (a - 0.499) / 0.499), al / (1.5 * R), be / (1.5 * R)], axis=1)
    y = tracer.surrogate_predict(feat).cpu().numpy().astype(np.float64)
    cls = np.argmax(y[:, :3], axis=1)
    assert np.array_equal(cls, got["cls"][idx] & 7)
    esc = np.nonzero(cls == 0)[0]
    d = []
    for k in esc:
        e1, e2, e3 = frame64(v.r_cam, inc, v.phi_cam, al[k], be[k])
        N = y[k, 3] * e1 + y[k, 4] * e2 + y[k, 5] * e3
        N /= np.linalg.norm(N)
        t, p = got["theta_inf"][idx[k]], got["phi_inf"][idx[k]]
        M = np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])
        d.append(math.atan2(np.linalg.norm(np.cross(N, M)), np.dot(N, M)))
    d = np.array(d)
    disk = np.nonzero(cls == 2)[0]
    rr = np.abs(got["r_hit"][idx[disk]] / np.exp(y[disk, 6]) - 1.0)
    gr = np.abs(got["g"][idx[disk]] / np.exp(y[disk, 7]) - 1.0)
    q = lambda x: np.quantile(x, [0.5, 0.99, 1.0])
    print(f"{len(esc)} escaped, {len(disk)} disk answered; p50 p99 max: sky {q(d)} rad, r_hit {q(rr)}, g {q(gr)}",
          flush=True)
    # inside the trained square the disk (to 20 M) fills most of the plane: escaped rays
    # answered by the network are a few hundred at most here
    assert len(esc) > 50 and len(disk) > 100
    for x in (d, rr, gr):
        assert np.median(x) < 1e-5 and np.quantile(x, 0.99) < 2e-3 and x.max() < 5e-2


def _domain_view(**kw):
    """a small Bardeen view of the network's scene (a = 0.9, r = 40 M, i = 60 deg), with
    fields replaced by `kw`"""
    base = dict(width=48, height=32, spin=0.9, r_cam=40.0, theta_cam=math.radians(60.0), phi_cam=0.3,
                half_width=7.0, half_height=4.7, disk_inner=0.0, disk_outer=20.0, disk_prograde=1, disk_enabled=1,
                escape_radius=1000.0, capture_factor=1.01)
    base.update(kw)
    return bsa.View(**base)


@pytest.mark.parametrize("change", [
    dict(disk_enabled=0), dict(disk_prograde=0), dict(disk_outer=30.0), dict(disk_inner=3.0),
    dict(theta_cam=math.radians(120.0)), dict(theta_cam=math.radians(2.0)), dict(r_cam=8.0), dict(r_cam=2000.0),
    dict(spin=0.999), dict(spin=-0.5), dict(capture_factor=1.05), dict(escape_radius=500.0),
    dict(camera_kind=bsa.CAMERA_PINHOLE),
])
def test_views_outside_the_domain_go_to_the_integrator(tracer, change):
    """The shipped network covers Bardeen planes of the scene it was trained on (prograde disk
    from the ISCO to 20 M, capture 1.01 r+, escape 1000 M) with a in [0, 0.998], i in
    [5, 85] deg and r_cam in [10, 1000].  A view outside that, traced with use_surrogate = 1
    beside a covered view in the same call: none of its rays is answered, and every one is
    the mixed integrator's own result; the covered view is still answered."""
    from black_simulation_amd.train_surrogate import load_params

    tracer.surrogate_load(load_params())
    inside, outside = _domain_view(), _domain_view(**change)
    opts = bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED)
    got = bsa.to_numpy(tracer.trace_views([inside, outside], opts))
    want = bsa.to_numpy(tracer.trace_views([outside], bsa.TraceOpts(precision=bsa.PREC_MIXED)))
    n = inside.rays
    assert ((got["cls"][:n] & bsa.FLAG_SURROGATE) != 0).mean() > 0.5
    assert not ((got["cls"][n:] & bsa.FLAG_SURROGATE) != 0).any()
    for k in ("cls", "steps", "theta_inf", "phi_inf", "r_hit", "g"):
        assert np.array_equal(got[k][n:], want[k], equal_nan=True), k


def test_rays_beyond_the_trained_image_plane_go_to_the_integrator(tracer):
    """Training draws alpha and beta within +-1.5 mean shadow radii: on a plane twice as wide
    the rays outside that square are all traced by the integrator.  Narrowing the domain's
    extent narrows the answered square with it; an extent beyond 1.5 is refused."""
    from black_simulation_amd.train_surrogate import load_params

    tracer.surrogate_load(load_params())
    default = tracer.surrogate_domain()
    R = tracer.surrogate_shadow_radius(0.9, math.radians(60.0))
    v = _domain_view(width=64, height=64, half_width=3.0 * R, half_height=3.0 * R)
    idx = np.arange(v.rays)
    al = -v.half_width + (idx % v.width + 0.5) * (2 * v.half_width / v.width)
    be = v.half_height - (idx // v.width + 0.5) * (2 * v.half_height / v.height)
    try:
        for extent in (1.5, 1.0):
            tracer.surrogate_set_domain(extent=extent)
            got = bsa.to_numpy(tracer.trace_views([v], bsa.TraceOpts(use_surrogate=1, precision=bsa.PREC_MIXED)))
            sur = (got["cls"] & bsa.FLAG_SURROGATE) != 0
            inside = (np.abs(al) <= extent * R * (1 - 1e-6)) & (np.abs(be) <= extent * R * (1 - 1e-6))
            far = (np.abs(al) > extent * R * (1 + 1e-6)) | (np.abs(be) > extent * R * (1 + 1e-6))
            assert sur[inside].mean() > 0.5, extent
            assert not sur[far].any(), extent
        with pytest.raises(RuntimeError):
            tracer.surrogate_set_domain(extent=2.0)
    finally:
        tracer.surrogate_set_domain(**default)
    assert tracer.surrogate_domain() == default
