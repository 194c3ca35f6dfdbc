"""The FP64 CPU oracle of the learned denoiser, pinned by finite differences
of its own loss (the network has no counterpart in the reference)."""
import numpy as np

import oracle


def _data(rng, b, h, w):
    x = np.empty((b, 10, h, w), np.float32)
    x[:, 0:3] = rng.gamma(1.0, 0.5, (b, 3, h, w))
    x[:, 3:6] = rng.uniform(0.2, 0.8, (b, 3, h, w))
    n = rng.normal(size=(b, 3, h, w))
    x[:, 6:9] = n / np.linalg.norm(n, axis=1, keepdims=True)
    x[:, 9] = rng.uniform(1, 10, (b, h, w))
    t = rng.gamma(2.0, 0.3, (b, h * w, 3)).astype(np.float32)
    return x, t


def _params(rng):
    p = np.zeros(oracle.cnn_num_params())
    for (cin, cout, taps), (ws, bs) in zip(oracle.CNN_LAYERS, oracle.cnn_layer_slices()):
        p[ws] = rng.normal(0, np.sqrt(2.0 / (taps * cin)), ws.stop - ws.start)
        p[bs] = rng.normal(0, 0.05, cout)
    return p


def test_parameter_count():
    assert oracle.cnn_num_params() == 474361


def test_softmax_kernel_is_a_convex_combination():
    rng = np.random.default_rng(0)
    x, _ = _data(rng, 1, 8, 8)
    p = _params(rng)
    out = oracle.cnn_forward(p, x[0], 8, 8)
    rad = x[0, :3].reshape(3, -1).T
    assert (out >= rad.min(axis=0) - 1e-12).all() and (out <= rad.max(axis=0) + 1e-12).all()


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    b, h, w = 2, 8, 8
    x, t = _data(rng, b, h, w)
    p = _params(rng)
    loss, g = oracle.cnn_loss_grad(p, x, t, b, h, w)
    assert loss > 0
    picks = []
    for ws, bs in oracle.cnn_layer_slices():
        picks += list(rng.integers(ws.start, ws.stop, 3)) + [bs.start]
    eps = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (oracle.cnn_loss_grad(pp, x, t, b, h, w, grads=False)[0] -
              oracle.cnn_loss_grad(pm, x, t, b, h, w, grads=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-4), (i, fd, g[i])


def test_device_precision_oracle_is_the_fp64_one_rounded():
    """The device-precision forward and gradients differ from the FP64 ones
    by bfloat16 rounding only: small but not zero, and equal where rounding
    changes nothing (weights and activations already bfloat16: an all-zero
    input gives the same output either way)."""
    rng = np.random.default_rng(2)
    h = w = 8
    # He-scaled weights, so that the softmax does not saturate (a saturated
    # one leaves only rounding noise in the gradients)
    p = np.zeros(oracle.cnn_num_params())
    for (cin, cout, taps), (ws, bs) in zip(oracle.CNN_LAYERS, oracle.cnn_layer_slices()):
        p[ws] = rng.normal(0, np.sqrt(2.0 / (taps * cin)), ws.stop - ws.start)
        p[bs] = rng.normal(0, 0.05, cout)
    x = rng.uniform(0.1, 1.0, (10, h, w)).astype(np.float32)
    a = oracle.cnn_forward(p, x, h, w)
    b = oracle.cnn_forward(p, x, h, w, bf16=True)
    rel = np.abs(a - b) / (np.abs(a) + 1e-2)
    assert 0 < rel.max() < 0.05
    t = rng.uniform(0.1, 1.0, (1, h * w, 3)).astype(np.float32)
    la, ga = oracle.cnn_loss_grad(p, x[None], t, 1, h, w)
    lb, gb = oracle.cnn_loss_grad(p, x[None], t, 1, h, w, bf16=True)
    assert abs(la - lb) < 0.05 * la
    cos = ga @ gb / (np.linalg.norm(ga) * np.linalg.norm(gb))
    assert 0.99 < cos < 1.0
