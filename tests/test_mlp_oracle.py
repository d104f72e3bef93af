"""The FP64 numpy surrogate network (oracle/mlp_oracle.py) against central finite
differences: its backward pass is what the GPU kernels' gradients are checked against."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def small_problem(seed=1, B=16):
    sizes = (5, 12, 12, 12, 12, 8)
    params = M.init_params(seed, sizes)
    params = [(W, np.random.default_rng(seed + l).normal(0, 0.1, b.shape)) for l, (W, b) in enumerate(params)]
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B, 5))
    cls = rng.integers(0, 3, B)
    tgt = rng.uniform(-1, 1, (B, 5))
    return params, x, cls, tgt


@pytest.mark.parametrize("act", ["silu", "hard"])
def test_backward_matches_finite_differences(act):
    params, x, cls, tgt = small_problem()
    y, cache = M.forward(params, x, act)
    _, gy = M.loss_and_grad_out(y, cls, tgt, w_sky=2.0, w_disk=0.5)
    grads = M.backward(params, cache, gy, act)
    # the hard SiLU's derivative jumps at +-3: keep every pre-activation clear of the kinks
    assert all(np.abs(np.abs(z) - 3.0).min() > 1e-4 for _, z in cache[:-1])
    rng = np.random.default_rng(0)
    h = 1e-6
    for l, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        for _ in range(6):
            i, j = rng.integers(0, W.shape[0]), rng.integers(0, W.shape[1])
            Wp, Wm = W.copy(), W.copy()
            Wp[i, j] += h
            Wm[i, j] -= h
            pp = [(Wp if k == l else p[0], p[1]) for k, p in enumerate(params)]
            pm = [(Wm if k == l else p[0], p[1]) for k, p in enumerate(params)]
            fd = (M.loss_fn(pp, x, cls, tgt, act, w_sky=2.0, w_disk=0.5)
                  - M.loss_fn(pm, x, cls, tgt, act, w_sky=2.0, w_disk=0.5)) / (2 * h)
            assert abs(fd - gW[i, j]) < 1e-7 + 1e-5 * abs(fd), (l, i, j, fd, gW[i, j])
        j = rng.integers(0, b.shape[0])
        bp, bm = b.copy(), b.copy()
        bp[j] += h
        bm[j] -= h
        pp = [(p[0], bp if k == l else p[1]) for k, p in enumerate(params)]
        pm = [(p[0], bm if k == l else p[1]) for k, p in enumerate(params)]
        fd = (M.loss_fn(pp, x, cls, tgt, act, w_sky=2.0, w_disk=0.5)
              - M.loss_fn(pm, x, cls, tgt, act, w_sky=2.0, w_disk=0.5)) / (2 * h)
        assert abs(fd - gb[j]) < 1e-7 + 1e-5 * abs(fd)


def test_adam_reduces_loss():
    params, x, cls, tgt = small_problem(B=64)
    st = M.adam_state(params)
    first = M.loss_fn(params, x, cls, tgt)
    for _ in range(200):
        y, cache = M.forward(params, x)
        _, gy = M.loss_and_grad_out(y, cls, tgt)
        params = M.adam_step(params, M.backward(params, cache, gy), st, 1e-2)
    assert M.loss_fn(params, x, cls, tgt) < 0.2 * first


def test_adam_first_step_is_lr_times_sign():
    """With bias correction the first Adam step moves every parameter by -lr sign(g)
    (up to eps); without it the step would be sqrt(1 - beta2) / (1 - beta1) = 0.316 times
    larger than lr in size."""
    params, x, cls, tgt = small_problem(B=32)
    y, cache = M.forward(params, x)
    _, gy = M.loss_and_grad_out(y, cls, tgt)
    grads = M.backward(params, cache, gy)
    new = M.adam_step(params, grads, M.adam_state(params), 1e-3)
    for (W, b), (gW, gb), (W1, b1) in zip(params, grads, new):
        for p, g, p1 in ((W, gW, W1), (b, gb, b1)):
            big = np.abs(g) > 1e-4
            np.testing.assert_allclose((p1 - p)[big], -1e-3 * np.sign(g[big]), rtol=1e-3)


def test_loss_closed_form():
    """Zero outputs: cross-entropy log 3 per ray, plus the squared targets of the class's
    regression outputs (sky for escaped, disk for disk rays) times their weights."""
    B = 3
    y = np.zeros((B, 8))
    cls = np.array([0, 1, 2])
    tgt = np.array([[0.5, -0.5, 0.25, 0.0, 0.0], [0.3, 0.3, 0.3, 0.3, 0.3], [0.0, 0.0, 0.0, 2.0, -1.0]])
    loss, _ = M.loss_and_grad_out(y, cls, tgt, w_sky=2.0, w_disk=0.5)
    want = (3 * np.log(3.0) + 2.0 * (0.25 + 0.25 + 0.0625) + 0.5 * (4.0 + 1.0)) / B
    assert loss == np.float64(want) or abs(loss - want) < 1e-12


def test_hard_silu_closed_forms():
    """z clamp(z/6 + 1/2, 0, 1): zero below -3, the identity above 3, z (z + 3) / 6 between
    (continuous at both ends), with derivative (2z + 3) / 6 there; SiLU's odd part is the
    same: f(z) - f(-z) = z."""
    z = np.array([-7.0, -3.0, -1.5, 0.0, 1.5, 3.0, 7.0])
    np.testing.assert_allclose(M.hsilu(z), [0.0, 0.0, -0.375, 0.0, 1.125, 3.0, 7.0], atol=1e-15)
    np.testing.assert_allclose(M.dhsilu(np.array([-4.0, -1.5, 0.0, 1.5, 4.0])), [0.0, 0.0, 0.5, 1.0, 1.0])
    zz = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(M.hsilu(zz) - M.hsilu(-zz), zz, atol=1e-14)
    np.testing.assert_allclose(M.silu(zz) - M.silu(-zz), zz, atol=1e-14)
