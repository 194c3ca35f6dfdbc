"""The learned denoiser's fallback to the reference's a-trous filter
(Source/Core/Scene.cpp:210-222 calls Util::atrous_denoise there): no network,
a network that refuses the image, and a network whose output is not finite
(NaN parameters) all give exactly the a-trous output, on the single-image
entry point and per view of the batched render + denoise launch; a working
network gives exactly its own output."""
import numpy as np
import pytest

import pico_amd as pa

pytestmark = pytest.mark.gpu

W, H = 72, 56


@pytest.fixture(scope="module")
def frame(ctx):
    scene = pa.Scene.cornell(1, pa.MICROFACET_GGX)
    ctx.upload_scene(scene)
    cam = scene.camera(W / H)
    n = W * H
    bufs = {k: ctx.alloc(12 * n) for k in ("pix", "var", "n", "p", "a", "s", "o", "ref")}
    bufs["cnt"] = ctx.alloc(4 * n)
    bufs["flag"] = ctx.alloc(4)
    ctx.render_d(cam, pa.params(W, H, 4, 5, frame_seed=2), bufs["pix"], bufs["cnt"], bufs["var"])
    ctx.gbuffer_d(cam, W, H, bufs["n"], bufs["p"], bufs["a"])
    ctx.atrous_d(bufs["pix"], bufs["n"], bufs["p"], bufs["a"], W, H, bufs["s"], bufs["ref"], tile=(16, 16))
    ctx.sync()
    yield scene, cam, bufs
    for b in bufs.values():
        b.free()


def _same(a, b):
    return np.array_equal(np.nan_to_num(a, nan=-7.0), np.nan_to_num(b, nan=-7.0)) and \
        np.array_equal(np.isnan(a), np.isnan(b))


def _run(ctx, net, bufs):
    bufs["flag"].upload(np.array([7], np.uint32))
    ctx.denoise_gbuffer_d(net, bufs["pix"], bufs["n"], bufs["p"], bufs["a"], W, H, bufs["s"], bufs["o"], bufs["flag"],
                          tile=(16, 16))
    ctx.sync()
    return bufs["o"].download(np.float32, 3 * W * H), int(bufs["flag"].download(np.uint32, 1)[0])


def test_no_network_gives_atrous(ctx, frame):
    _, _, bufs = frame
    out, flag = _run(ctx, None, bufs)
    assert flag == 1
    assert _same(out, bufs["ref"].download(np.float32, 3 * W * H))


def test_nan_network_gives_atrous(ctx, frame):
    _, _, bufs = frame
    net = pa.CNN(ctx, W, H, 1)
    net.set_params(np.full(net.n_params, np.nan, np.float32))
    out, flag = _run(ctx, net, bufs)
    net.close()
    assert flag == 1
    assert _same(out, bufs["ref"].download(np.float32, 3 * W * H))


def test_network_too_small_gives_atrous(ctx, frame):
    _, _, bufs = frame
    net = pa.CNN(ctx, 32, 32, 1)
    net.load()
    out, flag = _run(ctx, net, bufs)
    net.close()
    assert flag == 1
    assert _same(out, bufs["ref"].download(np.float32, 3 * W * H))


def test_working_network_keeps_its_output(ctx, frame):
    _, _, bufs = frame
    net = pa.CNN(ctx, W, H, 1)
    net.load()
    out, flag = _run(ctx, net, bufs)
    net.denoise_gbuffer_d(bufs["pix"], bufs["n"], bufs["p"], bufs["a"], W, H, bufs["ref"])
    ctx.sync()
    want = bufs["ref"].download(np.float32, 3 * W * H)
    net.close()
    assert flag == 0 and np.isfinite(out).all()
    assert np.array_equal(out, want)


def test_host_entry_point_reports_the_fallback(ctx, frame):
    scene, cam, bufs = frame
    n = W * H
    rgb = bufs["pix"].download(np.float32, 3 * n).reshape(n, 3)
    nn, pp, aa = (bufs[k].download(np.float32, 3 * n).reshape(n, 3) for k in ("n", "p", "a"))
    want = ctx.atrous(rgb, nn, pp, aa, W, H, (16, 16), 5)
    out, used = ctx.denoise_gbuffer(None, rgb, nn, pp, aa, W, H, tile=(16, 16))
    assert used and _same(out, want)
    net = pa.CNN(ctx, W, H, 1)
    net.load()
    out2, used2 = ctx.denoise_gbuffer(net, rgb, nn, pp, aa, W, H, tile=(16, 16))
    net.close()
    assert not used2 and np.isfinite(out2).all()


@pytest.mark.parametrize("flags", [0, pa.TUNE_SIDE_STREAM])
def test_batched_views_fall_back_per_view(ctx, flags):
    scene = pa.Scene.instanced(7, 4, 48, 200, 2000, 12.0, pa.MICROFACET_GGX)
    ctx.upload_scene(scene)
    specs = [(64, 48, 6, 5, 16), (40, 40, 4, 5, 8), (96, 32, 5, 4, 32)]
    old = ctx.get_tuning()
    ctx.set_tuning(old[0], old[1:3], flags)
    try:
        net = pa.CNN(ctx, 96, 96, 1)
        net.set_params(np.full(net.n_params, np.nan, np.float32))
        views, bufs, outs = [], [], []
        for vid, (w, h, spp, depth, tile) in enumerate(specs):
            n = w * h
            b = (ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n))
            v = pa.View()
            v.camera = scene.camera(w / h)
            v.params = pa.params(w, h, spp, depth, tile=tile, frame_seed=9, view_id=vid)
            v.d_pixels, v.d_sample_count, v.d_variance = b[0].ptr.value, b[1].ptr.value, b[2].ptr.value
            views.append(v)
            bufs.append(b)
            outs.append(ctx.alloc(12 * n))
        st = ctx.render_views_denoise_d(views, net, outs)
        assert [s.denoise_fallback for s in st] == [1, 1, 1]
        for v, b, o, (w, h, spp, depth, tile) in zip(views, bufs, outs, specs):
            n = w * h
            gb = [ctx.alloc(12 * n) for _ in range(5)]
            ctx.gbuffer_d(v.camera, w, h, gb[0], gb[1], gb[2])
            ctx.atrous_d(b[0], gb[0], gb[1], gb[2], w, h, gb[3], gb[4], tile=(tile, tile))
            ctx.sync()
            assert _same(o.download(np.float32, 3 * n), gb[4].download(np.float32, 3 * n)), (w, h)
        net.load()
        st = ctx.render_views_denoise_d(views, net, outs)
        assert [s.denoise_fallback for s in st] == [0, 0, 0]
        net.close()
    finally:
        ctx.set_tuning(old[0], old[1:3], old[3])
