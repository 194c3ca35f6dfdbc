"""Edge cases of the device path against the oracle: degenerate and ragged
image sizes, depth 0, a scene whose light list is empty, zero samples, empty
ray batches, filters on images smaller than a tile, and the learned filter on
sizes that are not a multiple of its pooling."""
import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cornell():
    return pa.Scene.cornell(1, pa.MICROFACET_GGX)


class _NoLights:
    """The Cornell box with an empty light list (its emitter stays visible)."""

    def __init__(self, scene):
        self.scene = scene  # keeps the arrays alive
        self.desc = pa.SceneDesc.from_buffer_copy(scene.desc)
        self.desc.n_lights = 0

    def camera(self, aspect):
        return self.scene.camera(aspect)


def _compare_render(ctx, scene, w, h, spp, depth, seed=3):
    ctx.upload_scene(scene)
    cam = scene.camera(w / h)
    p = pa.params(w, h, spp, depth, frame_seed=seed)
    gp, gc, gv, _ = ctx.render(cam, p)
    cp, cc, cv, _ = oracle.CpuScene(scene).render(cam, p, rng=oracle.RNG_PCG)
    assert np.array_equal(gc, cc)
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.98, (rel < 1e-3).mean()
    return gp, gc


@pytest.mark.parametrize("w,h", [(1, 1), (3, 1), (1, 5), (17, 29)])
def test_ragged_image_sizes(ctx, cornell, w, h):
    _compare_render(ctx, cornell, w, h, 8, 5)


def test_depth_zero(ctx, cornell):
    gp, _ = _compare_render(ctx, cornell, 16, 16, 4, 0)
    # depth 0: only camera-ray misses see anything (the box is closed but for
    # its open front, which the camera does not see)
    assert np.isfinite(gp).all()


def test_scene_without_lights(ctx, cornell):
    _compare_render(ctx, _NoLights(cornell), 24, 24, 8, 4)


def test_zero_samples_leave_buffers_untouched(ctx, cornell):
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    n = 8 * 8
    pix, cnt, var = ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n)
    marker = np.full(3 * n, 7.0, np.float32)
    pix.upload(marker)
    cnt.upload(np.full(n, 3, np.uint32))
    var.upload(marker)
    st = ctx.render_d(cam, pa.params(8, 8, 0, 5), pix, cnt, var)
    assert st.samples == 0
    assert np.array_equal(pix.download(np.float32, 3 * n), marker)
    assert np.array_equal(cnt.download(np.uint32, n), np.full(n, 3, np.uint32))


def test_empty_ray_batch(ctx, cornell):
    ctx.upload_scene(cornell)
    t, inst, prim, uv = ctx.closest_hit(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert t.size == 0 and inst.size == 0


@pytest.mark.parametrize("w,h,tile", [(1, 1, (16, 16)), (5, 3, (64, 64)), (9, 40, (16, 16))])
def test_atrous_small_images(ctx, w, h, tile):
    rng = np.random.default_rng(w * h)
    rgb = rng.gamma(2.0, 0.5, (w * h, 3)).astype(np.float32)
    n = rng.normal(size=(w * h, 3)).astype(np.float32)
    pos = (rng.normal(size=(w * h, 3)) * 3).astype(np.float32)
    alb = rng.uniform(0, 1, (w * h, 3)).astype(np.float32)
    g = ctx.atrous(rgb, n, pos, alb, w, h, tile, 5)
    c = oracle.atrous(rgb, n, pos, alb, w, h, tile, 5)
    both_nan = np.isnan(g) & np.isnan(c)
    np.testing.assert_allclose(np.where(both_nan, 0, g), np.where(both_nan, 0, c), rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("w,h", [(17, 29), (5, 3)])
def test_learned_filter_on_ragged_sizes(ctx, w, h):
    rng = np.random.default_rng(w + 100 * h)
    x = np.empty((10, h, w), np.float32)
    x[0:3] = rng.gamma(1.0, 0.5, (3, h, w))
    x[3:6] = rng.uniform(0.2, 0.8, (3, h, w))
    nn = rng.normal(size=(3, h, w))
    x[6:9] = nn / np.linalg.norm(nn, axis=0, keepdims=True)
    x[9] = rng.uniform(1, 10, (h, w))
    p = np.load(pa.weights_path(), allow_pickle=False)
    net = pa.CNN(ctx, w, h, 1)
    net.set_params(p)
    g = net.denoise(x, w, h)
    c = oracle.cnn_forward(p.astype(np.float64), x, h, w)
    err = np.abs(g - c) / (np.abs(c) + 1e-2)
    assert np.median(err) < 5e-3 and err.max() < 0.1, (np.median(err), err.max())


def test_zero_size_relmse_and_tonemap(ctx):
    d = ctx.alloc(64)
    assert ctx.relmse_d(d, d, 4.0, 0) == 0.0
    out = ctx.tonemap(np.zeros((0, 3), np.float32), 0, 5)
    assert out.size == 0


def test_relmse_size_limit_is_refused(ctx):
    d = ctx.alloc(64)
    with pytest.raises(pa.PicoError) as e:
        ctx.relmse_d(d, d, 4.0, (1 << 28) + 1)
    assert e.value.code == pa.PICO_ERR_LIMIT
