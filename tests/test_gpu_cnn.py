"""The learned denoiser on the device (BF16 MFMA, FP32 accumulation) against
its FP64 CPU oracle: forward output, loss, every parameter gradient, and a
training run that lowers the loss."""
import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


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
    p = np.zeros(oracle.cnn_num_params(), np.float32)
    for (cin, cout, taps), (ws, bs) in zip(oracle.CNN_LAYERS, oracle.cnn_layer_slices()):
        p[ws] = rng.normal(0, np.sqrt(2.0 / (taps * cin)), ws.stop - ws.start)
        p[bs] = rng.normal(0, 0.05, cout)
    # a head with some spread so the predicted kernels are not uniform
    # (logit std about 5); a much sharper softmax turns the BF16 rounding of
    # the logits (about 2^-9 relative) into weight errors of several percent
    ws, bs = oracle.cnn_layer_slices()[-1]
    p[ws] *= 5
    return p


@pytest.mark.parametrize("h,w", [(16, 16), (24, 40)])
def test_forward_matches_oracle(ctx, h, w):
    rng = np.random.default_rng(h + w)
    x, _ = _data(rng, 1, h, w)
    p = _params(rng)
    net = pa.CNN(ctx, w, h, 1)
    net.set_params(p)
    g = net.denoise(x[0], w, h)
    c = oracle.cnn_forward(p.astype(np.float64), x[0], h, w)
    err = np.abs(g - c) / (np.abs(c) + 1e-2)
    # bf16 activations and weights: ~3 significant digits per layer
    assert np.median(err) < 5e-3 and err.max() < 0.1, (np.median(err), err.max())


def _round_bf16(a):
    """float32 -> nearest bfloat16 (ties to even), back in float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("h,w,shipped", [(16, 16, False), (24, 40, False), (48, 64, True)])
def test_forward_matches_the_oracle_at_device_precision(ctx, h, w, shipped):
    """The oracle with the device's storage precision (bf16 weights, input and
    activations, float logits; FP64 sums): what is left is the FP32
    accumulation order, so the bound is an order of magnitude tighter than
    against the full-FP64 oracle."""
    rng = np.random.default_rng(10 * h + w)
    x, _ = _data(rng, 1, h, w)
    p = np.load(pa.weights_path(), allow_pickle=False) if shipped else _params(rng)
    net = pa.CNN(ctx, w, h, 1)
    net.set_params(p)
    g = net.denoise(x[0], w, h)
    c = oracle.cnn_forward(p.astype(np.float64), x[0], h, w, bf16=True)
    err = np.abs(g - c) / (np.abs(c) + 1e-2)
    print(f"[bf16 oracle] {h}x{w} median {np.median(err):.2e} p99 {np.percentile(err, 99):.2e} max {err.max():.2e}")
    assert np.median(err) < 1e-5 and np.percentile(err, 99) < 1e-2 and err.max() < 3e-2, \
        (np.median(err), np.percentile(err, 99), err.max())


def test_gradients_match_oracle(ctx):
    rng = np.random.default_rng(3)
    b, h, w = 2, 16, 16
    x, t = _data(rng, b, h, w)
    p = _params(rng)
    net = pa.CNN(ctx, w, h, b)
    net.set_params(p)
    gl, gg = net.loss_grad(x, t, b, w, h)
    cl, cg = oracle.cnn_loss_grad(p.astype(np.float64), x, t, b, h, w)
    assert abs(gl - cl) <= 2e-2 * cl
    for i, (ws, bs) in enumerate(oracle.cnn_layer_slices()):
        for s in (ws, bs):
            a, r = gg[s].astype(np.float64), cg[s]
            cos = a @ r / (np.linalg.norm(a) * np.linalg.norm(r) + 1e-30)
            rel = np.linalg.norm(a - r) / (np.linalg.norm(r) + 1e-30)
            assert cos > 0.99 and rel < 0.15, (i, s, cos, rel)
    # against the oracle at the device's storage precision (bf16 weights,
    # activations and stored gradients): only the FP32 sums differ
    bl, bg = oracle.cnn_loss_grad(p.astype(np.float64), x, t, b, h, w, bf16=True)
    worst = 0.0
    for i, (ws, bs) in enumerate(oracle.cnn_layer_slices()):
        for s in (ws, bs):
            a, r = gg[s].astype(np.float64), bg[s]
            rel = np.linalg.norm(a - r) / (np.linalg.norm(r) + 1e-30)
            worst = max(worst, rel)
    print(f"[bf16 oracle] loss rel {abs(gl - bl) / bl:.2e}, worst per-tensor gradient rel {worst:.2e}")
    assert abs(gl - bl) <= 1e-6 * bl and worst < 1e-2, (abs(gl - bl) / bl, worst)


def test_training_lowers_the_loss(ctx):
    rng = np.random.default_rng(4)
    b, h, w = 4, 32, 32
    x, _ = _data(rng, b, h, w)
    # target: a smoothed version of the noisy radiance (a denoising task)
    rad = x[:, 0:3]
    sm = rad.copy()
    for ax in (2, 3):
        sm = (np.roll(sm, 1, ax) + sm + np.roll(sm, -1, ax)) / 3
    t = np.ascontiguousarray(sm.reshape(b, 3, h * w).transpose(0, 2, 1))
    net = pa.CNN(ctx, w, h, b)
    net.init(7)
    d_x, d_t = ctx.alloc(x.nbytes), ctx.alloc(t.nbytes)
    d_x.upload(x)
    d_t.upload(t)
    losses = [net.train_step_d(d_x, d_t, b, w, h, 2e-3) for _ in range(60)]
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])


def test_gbuffer_entry_point_equals_planar_inputs(ctx):
    """The G-buffer form of the filter assembles radiance, albedo, normal and
    |position| into the planar input and gives the planar form's output."""
    rng = np.random.default_rng(5)
    w, h = 40, 24
    rgb = rng.gamma(1.0, 0.5, (w * h, 3)).astype(np.float32)
    nrm = rng.normal(size=(w * h, 3)).astype(np.float32)
    pos = rng.normal(0, 20, (w * h, 3)).astype(np.float32)
    alb = rng.uniform(0.2, 0.8, (w * h, 3)).astype(np.float32)
    x = np.concatenate([rgb.T, alb.T, nrm.T, np.sqrt((pos.astype(np.float64) ** 2).sum(1))[None].astype(np.float32)])
    net = pa.CNN(ctx, w, h, 1)
    net.set_params(_params(rng))
    b = net.denoise_gbuffer(rgb, nrm, pos, alb, w, h)
    d = [ctx.alloc(12 * w * h) for _ in range(4)]
    for arr, src in zip(d, (rgb, nrm, pos, alb)):
        arr.upload(src)
    d_in = ctx.alloc(40 * w * h)
    ctx.cnn_inputs_d(*d, w, h, d_in)
    xi = d_in.download(np.float32, 10 * w * h).reshape(10, w * h)
    np.testing.assert_array_equal(xi[:9], x[:9])
    np.testing.assert_allclose(xi[9], x[9], rtol=1e-6)
    np.testing.assert_array_equal(b, net.denoise(xi, w, h))


def test_trained_network_gains_over_the_methods_without_it(ctx):
    """Held-out procedural scenes (seeds outside the training range), 4-spp
    input, 512-spp FP64 target: the shipped trained network beats the raw
    input, the fixed-width a-trous filter and the untrained network."""
    from pico_amd.kpn_train import PairRenderer, relmse
    S = 128
    pr = PairRenderer(ctx, S, S, spp_in=4, spp_target=512, depth=8)
    trained, untrained = pa.CNN(ctx, S, S, 1), pa.CNN(ctx, S, S, 1)
    trained.load()
    untrained.init(1)
    d_x, d_t, d_o, d_s = ctx.alloc(40 * S * S), ctx.alloc(12 * S * S), ctx.alloc(12 * S * S), ctx.alloc(12 * S * S)
    err = {"raw": [], "atrous": [], "untrained": [], "trained": []}
    for seed in (20000, 20001, 20002, 20003, 20004, 20005):
        scene = pa.Scene.training(seed)
        pr.render(scene, seed, d_x, d_t)
        tgt = d_t.download(np.float32, 3 * S * S).reshape(-1, 3)
        rgb = np.ascontiguousarray(d_x.download(np.float32, 3 * S * S).reshape(3, -1).T)
        err["raw"].append(relmse(rgb, tgt))
        pr.d_pix.upload(rgb)
        ctx.gbuffer_d(scene.camera(1.0), S, S, pr.d_n, pr.d_p, pr.d_a)
        ctx.atrous_fixed_d(pr.d_pix, pr.d_n, pr.d_p, pr.d_a, S, S, d_s, d_o)
        err["atrous"].append(relmse(d_o.download(np.float32, 3 * S * S).reshape(-1, 3), tgt))
        for name, net in (("untrained", untrained), ("trained", trained)):
            net.denoise_d(d_x, S, S, d_o)
            err[name].append(relmse(d_o.download(np.float32, 3 * S * S).reshape(-1, 3), tgt))
    m = {k: float(np.mean(v)) for k, v in err.items()}
    print("held-out relMSE", m)
    assert m["trained"] < m["atrous"] and m["trained"] < m["raw"] and m["trained"] < m["untrained"], m
