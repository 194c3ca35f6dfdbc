"""Device renderer semantics against the oracle: accumulation and early-out,
statistical agreement with the reference's own generators, G-buffer, tone
mapping, FP64 accumulation, relMSE, batched views, resume, cancel, errors."""
import ctypes as C

import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cornell():
    return pa.Scene.cornell(1, pa.MICROFACET_GGX)


@pytest.fixture(scope="module")
def small_inst():
    return pa.Scene.instanced(7, 4, 48, 200, 2000, 12.0, pa.MICROFACET_GGX)


def test_render_accumulation_matches_oracle(ctx, cornell):
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    p = pa.params(64, 64, 10, 5, frame_seed=3)
    gp, gc, gv, st = ctx.render(cam, p)
    cp, cc, cv, cst = oracle.CpuScene(cornell).render(cam, p, rng=oracle.RNG_PCG)
    assert (gc == cc).mean() >= 0.999
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.99
    assert st.samples == cst.samples or abs(int(st.samples) - int(cst.samples)) <= 0.001 * cst.samples
    assert st.accumulated_rays > 0 and st.render_ms > 0


def test_early_out_on_constant_pixels(ctx, small_inst):
    """Pixels that see only the constant sky have zero variance after 6 samples
    and stop there (Source/Core/Scene.cpp:196)."""
    ctx.upload_scene(small_inst)
    cam = pa.Camera((6.0, 6.0, -3.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 1.0, 100.0)  # looking up, away
    p = pa.params(32, 32, 20, 4, frame_seed=1)
    _, gc, _, _ = ctx.render(cam, p)
    _, cc, _, _ = oracle.CpuScene(small_inst).render(cam, p, rng=oracle.RNG_PCG)
    assert np.array_equal(gc, cc)
    assert (gc == 6).mean() > 0.5


@pytest.mark.parametrize("which", ["cornell", "small_inst"])
def test_statistical_agreement_with_reference_generators(ctx, cornell, small_inst, which):
    """Per-pixel radiance over 16 frame seeds: device (PCG32 streams) vs the
    oracle with the reference's own generators (reproduces the compiled
    reference sample by sample).  Independent estimators of one expectation."""
    scene = cornell if which == "cornell" else small_inst
    ctx.upload_scene(scene)
    cam = scene.camera(1.0)
    w = h = 40
    cs = oracle.CpuScene(scene)
    g, c = [], []
    for seed in range(16):
        p = pa.params(w, h, 8, 5, frame_seed=100 + seed)
        g.append(ctx.render(cam, p)[0])
        c.append(cs.render(cam, p, rng=oracle.RNG_REFERENCE)[0])
    g, c = np.array(g, np.float64), np.array(c, np.float64)
    mg, mc = g.mean(0), c.mean(0)
    se = np.sqrt(g.var(0, ddof=1) / 16 + c.var(0, ddof=1) / 16)
    ok = np.abs(mg - mc) <= 3 * se + 1e-6
    assert ok.mean() >= 0.99, ok.mean()
    # whole image: means agree within 3 standard errors
    img_g, img_c = g.mean(axis=(1, 2)), c.mean(axis=(1, 2))
    se_img = np.sqrt(img_g.var(ddof=1) / 16 + img_c.var(ddof=1) / 16)
    assert abs(img_g.mean() - img_c.mean()) <= 3 * se_img


def test_gbuffer_matches_oracle(ctx, cornell, small_inst):
    """Position agrees everywhere; normal and albedo may differ only where two
    surfaces tie at the same point (pixel-corner rays on box corners)."""
    for scene in (cornell, small_inst):
        ctx.upload_scene(scene)
        cam = scene.camera(1.0)
        g = ctx.gbuffer(cam, 80, 60)
        c = oracle.CpuScene(scene).gbuffer(cam, 80, 60)
        pos_ok = np.isclose(g[1], c[1], rtol=1e-4, atol=1e-4).all(axis=1)
        assert pos_ok.all()
        for a, b in (g[0], c[0]), (g[2], c[2]):
            close = np.isclose(a, b, rtol=1e-4, atol=1e-4).all(axis=1)
            assert close.mean() >= 0.98


def test_tonemap_matches_oracle(ctx):
    rng = np.random.default_rng(4)
    rgb = rng.gamma(1.5, 2.0, (300 * 200, 3)).astype(np.float32)
    np.testing.assert_allclose(ctx.tonemap(rgb, 300, 200), oracle.tonemap(rgb, 300, 200), rtol=1e-6, atol=1e-7)


def test_fp64_accumulation_and_relmse(ctx, cornell):
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    w = h = 32
    n = w * h
    p = pa.params(w, h, 6, 5, frame_seed=9)
    d_sum = ctx.alloc(24 * n)
    d_sum.zero()
    ctx.render_accum64_d(cam, p, 0, d_sum)
    s64 = d_sum.download(np.float64, 3 * n).reshape(n, 3)
    samples = ctx.render_samples(cam, p, 0)
    np.testing.assert_allclose(s64, samples.astype(np.float64).sum(0), rtol=1e-6, atol=1e-6)
    img = (samples[3] + 0.01).astype(np.float32)
    d_img = ctx.alloc(12 * n)
    d_img.upload(img)
    got = ctx.relmse_d(d_img, d_sum, 6.0, n)
    want = oracle.relmse(img, s64, 6.0)
    assert abs(got - want) <= 1e-12 * max(1.0, want)


def test_batched_views_equal_single_views(ctx, cornell):
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    specs = [(24, 24, 3, 5), (40, 16, 7, 4), (8, 8, 12, 5)]
    views, bufs = [], []
    for vid, (w, h, spp, depth) in enumerate(specs):
        n = w * h
        b = (ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n))
        bufs.append(b)
        v = pa.View()
        v.camera = cam
        v.params = pa.params(w, h, spp, depth, frame_seed=5, view_id=vid)
        v.d_pixels, v.d_sample_count, v.d_variance = b[0].ptr.value, b[1].ptr.value, b[2].ptr.value
        views.append(v)
    stats = ctx.render_views_d(views)
    for vid, (w, h, spp, depth) in enumerate(specs):
        n = w * h
        pix = bufs[vid][0].download(np.float32, 3 * n).reshape(n, 3)
        cnt = bufs[vid][1].download(np.uint32, n)
        sp, sc, _, _ = ctx.render(cam, pa.params(w, h, spp, depth, frame_seed=5, view_id=vid))
        assert np.array_equal(cnt, sc)
        np.testing.assert_array_equal(pix, sp)
        assert stats[vid].samples == sc.sum() if spp <= 6 else stats[vid].samples > 0


def test_resume_continues_accumulation(ctx, cornell):
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    p = pa.params(24, 24, 3, 5, frame_seed=2)
    gp, gc, gv, _ = ctx.render(cam, p)
    p2 = pa.params(24, 24, 5, 5, frame_seed=2, flags=pa.RENDER_RESUME)
    gp, gc, gv, _ = ctx.render(cam, p2, gp, gc, gv)
    cs = oracle.CpuScene(cornell)
    cp, cc, cv, _ = cs.render(cam, p, rng=oracle.RNG_PCG)
    # the oracle continues from its own buffers the same way
    st = pa.Stats()
    oracle.cpu_lib().cpuref_render(cs.h, C.byref(cam), C.byref(p2), oracle.RNG_PCG, 0, oracle._p(cp), oracle._p(cc),
                                   oracle._p(cv), C.byref(st))
    assert np.array_equal(gc, cc)
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.99


def test_cancel_before_start(ctx, cornell):
    ctx.upload_scene(cornell)
    flag = C.c_int32(1)
    pix, cnt, var, st = ctx.render(cornell.camera(1.0), pa.params(16, 16, 4, 3), cancel=flag)
    assert st.cancelled == 1 and cnt.sum() == 0


def test_errors_are_reported(ctx, cornell):
    bad = pa.Scene.cornell(1)
    idx = np.ctypeslib.as_array(C.cast(bad.desc.indices, C.POINTER(C.c_uint32)), (3,))
    saved = idx[0]
    idx[0] = 1000  # index beyond the mesh's vertices
    with pytest.raises(pa.PicoError) as e:
        ctx.upload_scene(bad)
    idx[0] = saved
    assert e.value.code == pa.PICO_ERR_ARG and "index" in str(e.value)
    with pytest.raises(pa.PicoError):
        ctx.render(cornell.camera(1.0), pa.params(0, 16, 1, 1))
    with pytest.raises(pa.PicoError):
        ctx.render(cornell.camera(1.0), pa.params(16, 16, 1, 200))


@pytest.mark.parametrize("w,h,queue", [(160, 90, 8 * 160 * 90), (1920, 1080, 0)])
def test_multi_chunk_render_matches_oracle(ctx, w, h, queue):
    """Renders that run in several chunks of samples (6, then 8 at a time,
    with early-outed pixels compacted away between chunks) against the oracle
    with the identical PCG streams: the benchmark scene, at a reduced size
    with the queue depth forcing the 1080p chunking, and at 1080p itself."""
    scene = pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_GGX)
    ctx.upload_scene(scene)
    cam = scene.camera(w / h)
    old = ctx.get_tuning()
    try:
        if queue:
            ctx.set_tuning(queue_paths=queue)
        p = pa.params(w, h, 16, 8, frame_seed=7)
        gp, gc, gv, _ = ctx.render(cam, p)
    finally:
        ctx.set_tuning(old[0], old[1:3], old[3])
    cp, cc, cv, _ = oracle.CpuScene(scene).render(cam, p, rng=oracle.RNG_PCG, threads=16)
    assert (gc == cc).mean() >= 0.999, (gc == cc).mean()
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.99, (rel < 1e-3).mean()


def test_bounce_launches_follow_the_last_hand_over(ctx, small_inst, cornell):
    """A render launches the bounces of a chunk up to the point where an
    earlier render's chunk of the same shape (pixels, samples, depth) handed
    its paths over to the path-by-path kernel, and hands over there at the
    latest.  The output does not depend on that point: renders of the open
    instanced scene (it hands over after a few bounces) and of the closed
    Cornell box (its paths live longer, so the hint learnt on the other scene
    forces its hand-over with over a million paths left), alternated in one
    context, give the pixels of a fresh context bit for bit."""
    p = pa.params(1024, 1024, 2, 8, frame_seed=5)
    want = {}
    fresh = pa.Context(0)
    try:
        for name, sc in (("inst", small_inst), ("cornell", cornell)):
            fresh.upload_scene(sc)
            want[name] = fresh.render(sc.camera(1.0), p)
    finally:
        fresh.close()
    for name, sc, reps in (("inst", small_inst, 2), ("cornell", cornell, 2), ("inst", small_inst, 1)):
        ctx.upload_scene(sc)
        for _ in range(reps):
            gp, gc, gv, st = ctx.render(sc.camera(1.0), p)
            w = want[name]
            assert np.array_equal(gc, w[1]), name
            assert np.array_equal(gp.view(np.uint32), w[0].view(np.uint32)), name
            assert np.array_equal(gv.view(np.uint32), w[2].view(np.uint32)), name
            assert st.extension_rays == w[3].extension_rays and st.shadow_rays == w[3].shadow_rays, name
