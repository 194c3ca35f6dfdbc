"""Batched uneven views: the launch shape never changes the output, the
autotuner is deterministic once it has measured and never picks a shape
slower than the default on its own measurements, and the render + denoise
launch gives each view exactly what the separate calls give."""
import numpy as np
import pytest

import pico_amd as pa

pytestmark = pytest.mark.gpu

SPECS = [(96, 64, 9, 5), (40, 40, 3, 5), (128, 32, 16, 4), (8, 8, 12, 5), (64, 96, 7, 5)]


@pytest.fixture(scope="module")
def scene():
    return pa.Scene.instanced(7, 4, 48, 200, 2000, 12.0, pa.MICROFACET_GGX)


def _views(ctx, scene, specs, seed=5):
    views, bufs = [], []
    for vid, (w, h, spp, depth) in enumerate(specs):
        n = w * h
        b = (ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n))
        v = pa.View()
        v.camera = scene.camera(w / h)
        v.params = pa.params(w, h, spp, depth, frame_seed=seed, view_id=vid)
        v.d_pixels, v.d_sample_count, v.d_variance = b[0].ptr.value, b[1].ptr.value, b[2].ptr.value
        views.append(v)
        bufs.append(b)
    return views, bufs


def _read(bufs, specs):
    out = []
    for b, (w, h, _, _) in zip(bufs, specs):
        n = w * h
        out.append((b[0].download(np.float32, 3 * n), b[1].download(np.uint32, n), b[2].download(np.float32, 3 * n)))
    return out


@pytest.fixture
def restore_tuning(ctx):
    old = ctx.get_tuning()
    yield
    ctx.set_tuning(old[0], old[1:3], old[3])


def test_output_does_not_depend_on_the_launch_shape(ctx, scene, restore_tuning):
    ctx.upload_scene(scene)
    views, bufs = _views(ctx, scene, SPECS)
    results = []
    for q, tile, flags in [(16 << 20, (0, 0), 0), (4096, (8, 8), 0), (20000, (32, 4), 0), (1 << 20, (16, 16), 0),
                           (1 << 20, (8, 8), pa.TUNE_SHORTEST_FIRST)]:
        ctx.set_tuning(q, tile, flags)
        ctx.render_views_d(views)
        results.append(_read(bufs, SPECS))
    for r in results[1:]:
        for (p0, c0, v0), (p1, c1, v1) in zip(results[0], r):
            assert np.array_equal(c0, c1)
            assert np.array_equal(p0, p1)
            assert np.array_equal(v0, v1)


def test_per_view_completion_times(ctx, scene):
    ctx.upload_scene(scene)
    views, bufs = _views(ctx, scene, SPECS)
    stats = ctx.render_views_d(views)
    for st, (w, h, spp, _) in zip(stats, SPECS):
        assert 0 < st.render_ms <= st.done_ms + 1e-6
        assert st.samples > 0
    # the view with the most samples finishes last
    last = max(range(len(SPECS)), key=lambda i: SPECS[i][2])
    assert stats[last].render_ms >= max(s.render_ms for s in stats) - 1e-6


def test_autotuner_is_deterministic_and_never_worse_than_default(ctx, scene, restore_tuning):
    ctx.upload_scene(scene)
    views, bufs = _views(ctx, scene, SPECS)
    t1, c1, d1 = ctx.autotune_views(views, calib_spp=4)
    assert c1 <= d1
    assert ctx.get_tuning() == t1
    ctx.set_tuning()  # back to the default; the cached choice comes back without measuring
    t2, c2, d2 = ctx.autotune_views(views, calib_spp=4)
    assert (t2, c2, d2) == (t1, c1, d1)
    assert ctx.get_tuning() == t1


@pytest.mark.parametrize("flags", [0, pa.TUNE_SIDE_STREAM, pa.TUNE_SHORTEST_FIRST])
def test_render_and_denoise_launch_equals_separate_calls(ctx, scene, restore_tuning, flags):
    ctx.upload_scene(scene)
    ctx.set_tuning(flags=flags)
    specs = SPECS[:3]
    views, bufs = _views(ctx, scene, specs)
    net = pa.CNN(ctx, 128, 96, 1)
    net.init(3)
    outs = [ctx.alloc(12 * w * h) for (w, h, _, _) in specs]
    stats = ctx.render_views_denoise_d(views, net, outs)
    got = [o.download(np.float32, 3 * w * h) for o, (w, h, _, _) in zip(outs, specs)]
    rendered = _read(bufs, specs)
    for k, (w, h, spp, depth) in enumerate(specs):
        n = w * h
        dn, dp, da, dr, do = (ctx.alloc(12 * n) for _ in range(5))
        dr.upload(rendered[k][0])
        ctx.gbuffer_d(views[k].camera, w, h, dn, dp, da)
        net.denoise_gbuffer_d(dr, dn, dp, da, w, h, do)
        np.testing.assert_array_equal(got[k], do.download(np.float32, 3 * n))
        assert stats[k].done_ms >= stats[k].render_ms > 0


def test_shortest_first_finishes_cheap_views_first(ctx, scene, restore_tuning):
    ctx.upload_scene(scene)
    ctx.set_tuning(flags=pa.TUNE_SHORTEST_FIRST)
    views, bufs = _views(ctx, scene, SPECS)
    stats = ctx.render_views_d(views)
    work = [w * h * spp for (w, h, spp, _) in SPECS]
    cheap, dear = int(np.argmin(work)), int(np.argmax(work))
    assert stats[cheap].render_ms < stats[dear].render_ms


def test_host_buffer_views_equal_device_views(ctx, scene):
    """pico_hip_render_views (host buffers) gives what the device-buffer
    launch gives, view by view."""
    ctx.upload_scene(scene)
    specs = SPECS[:3]
    views, bufs = _views(ctx, scene, specs, seed=8)
    ctx.render_views_d(views)
    dev = _read(bufs, specs)
    outs, st = ctx.render_views([v.camera for v in views], [v.params for v in views])
    for (dp, dc, dvar), (hp, hc, hv) in zip(dev, outs):
        assert np.array_equal(dc, hc)
        assert np.array_equal(dp, hp.reshape(-1))
        assert np.array_equal(dvar, hv.reshape(-1))
    assert all(s.samples > 0 for s in st)


def test_progressive_snapshots(ctx, scene):
    """The pass callback sees growing sample counts after every chunk, the
    last snapshot is the final image, and a non-zero return stops the render
    with the samples so far kept (as a cancel)."""
    ctx.upload_scene(scene)
    w, h = 48, 40
    cam = scene.camera(w / h)
    p = pa.params(w, h, 24, 5, frame_seed=6)
    old = ctx.get_tuning()
    ctx.set_tuning(w * h * 4, old[1:3], old[3])  # four samples per chunk after the first
    try:
        seen = []
        pix, cnt, _, st = ctx.render_progressive(cam, p, lambda a, b, done: seen.append((a, b, done)) and False)
        ref, rcnt, _, _ = ctx.render(cam, p)
        assert len(seen) >= 3 and [s[2] for s in seen] == sorted(s[2] for s in seen) and seen[-1][2] == 24
        assert all((b1 >= b0).all() for (_, b0, _), (_, b1, _) in zip(seen, seen[1:]))
        assert np.array_equal(seen[-1][0], pix) and np.array_equal(pix, ref) and np.array_equal(cnt, rcnt)
        calls = []
        pix2, cnt2, _, st2 = ctx.render_progressive(cam, p, lambda a, b, done: calls.append(done) or len(calls) >= 2)
        assert st2.cancelled == 1 and len(calls) == 2 and cnt2.max() == calls[-1] < 24
    finally:
        ctx.set_tuning(old[0], old[1:3], old[3])
