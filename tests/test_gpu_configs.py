"""The benchmark configurations at their stated sizes, device against the
oracle (per-pixel PCG streams identical on both sides):

C1  Cornell box (scene seed 1), 256x256, 16 spp, depth 5, 16x16 tiles, GGX and
    the reference's Beckmann_All: the render path by path, the same pixels
    over frame seeds 0..15 against the reference's own generators (3 standard
    errors), and reference-mode a-trous on the device's own frame and
    G-buffer (Source/Core/Scene.cpp:143-235, Source/Util/Denoisers.cpp:18-126).
C2  instanced scene (scene seed 2, 1024 instances, ~8.3M triangles),
    1920x1080: primary-hit ids on every pixel, and the 64-spp, depth-8 render.
C3  the same scene at 4 and 16 spp: both a-trous modes on the device's own
    1080p frame and G-buffer, and relMSE against an FP64 accumulation.
C4  the shipped learned denoiser on the 1080p frame (every output pixel is a
    convex combination of its clamped 5x5 neighbourhood) and against the FP64
    network oracle at 480x270.
C5  the benchmark's 64 views (square 128..1024, 4..64 spp, view seed 3) in one
    launch, each bit-identical to its own single-view render, and the render +
    denoise launch equal to the separate calls."""
import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


def _same_or_both_nan(a, b, rtol=1e-4, atol=1e-6):
    nan = np.isnan(a) & np.isnan(b)
    assert (np.isnan(a) == np.isnan(b)).all(), (np.isnan(a).sum(), np.isnan(b).sum())
    np.testing.assert_allclose(np.where(nan, 0, a), np.where(nan, 0, b), rtol=rtol, atol=atol)


@pytest.fixture(scope="module")
def bench_scene():
    return pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_GGX)


# ---- C1 -------------------------------------------------------------------

@pytest.mark.parametrize("mf", [pa.MICROFACET_GGX, pa.MICROFACET_BECKMANN_ALL])
def test_c1_cornell_render_and_atrous(ctx, mf):
    scene = pa.Scene.cornell(1, mf)
    ctx.upload_scene(scene)
    cam = scene.camera(1.0)
    W = H = 256
    p = pa.params(W, H, 16, 5, tile=16, frame_seed=0)
    gp, gc, _, _ = ctx.render(cam, p)
    cp, cc, _, _ = oracle.CpuScene(scene).render(cam, p, rng=oracle.RNG_PCG)
    assert (gc == cc).mean() >= 0.999, (gc == cc).mean()
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.99, (rel < 1e-3).mean()
    # the device's G-buffer against the oracle's: positions everywhere
    n, pos, alb = ctx.gbuffer(cam, W, H)
    cn, cpos, calb = oracle.CpuScene(scene).gbuffer(cam, W, H)
    assert np.isclose(pos, cpos, rtol=1e-4, atol=1e-4).all()
    # reference-mode a-trous, 16x16 tiles, 5 levels, on the device's frame
    g = ctx.atrous(gp, n, pos, alb, W, H, (16, 16), 5)
    c = oracle.atrous(gp, n, pos, alb, W, H, (16, 16), 5)
    _same_or_both_nan(g, c)


def test_c1_cornell_statistics_over_frame_seeds(ctx):
    """Per-pixel radiance over frame seeds 0..15 at 256x256, 16 spp: the
    device's PCG streams against the reference's own generators (xorshift,
    Hammersley, mt19937_64), which the oracle reproduces sample by sample."""
    scene = pa.Scene.cornell(1, pa.MICROFACET_GGX)
    ctx.upload_scene(scene)
    cam = scene.camera(1.0)
    cs = oracle.CpuScene(scene)
    g, c = [], []
    for seed in range(16):
        p = pa.params(256, 256, 16, 5, tile=16, frame_seed=seed)
        g.append(ctx.render(cam, p)[0])
        c.append(cs.render(cam, p, rng=oracle.RNG_REFERENCE)[0])
    g, c = np.array(g, np.float64), np.array(c, np.float64)
    se = np.sqrt(g.var(0, ddof=1) / 16 + c.var(0, ddof=1) / 16)
    ok = np.abs(g.mean(0) - c.mean(0)) <= 3 * se + 1e-6
    assert ok.mean() >= 0.99, ok.mean()
    img_g, img_c = g.mean(axis=(1, 2)), c.mean(axis=(1, 2))
    se_img = np.sqrt(img_g.var(ddof=1) / 16 + img_c.var(ddof=1) / 16)
    assert abs(img_g.mean() - img_c.mean()) <= 3 * se_img


# ---- C2 -------------------------------------------------------------------

@pytest.mark.parametrize("xi", [(0.0, 0.0), (0.37, 0.61)])
def test_c2_primary_ids_1080p(ctx, bench_scene, xi):
    ctx.upload_scene(bench_scene)
    cam = bench_scene.camera(1920 / 1080)
    gi, gp = ctx.primary_ids(cam, 1920, 1080, xi)
    ci, cp = oracle.CpuScene(bench_scene).primary_ids(cam, 1920, 1080, xi)
    assert (ci != 0xffffffff).mean() > 0.3
    assert ((gi == ci) & (gp == cp)).mean() >= 0.9999


def test_c2_render_1080p_64spp(ctx, bench_scene):
    ctx.upload_scene(bench_scene)
    cam = bench_scene.camera(1920 / 1080)
    p = pa.params(1920, 1080, 64, 8, tile=16, frame_seed=1)
    gp, gc, _, st = ctx.render(cam, p)
    cp, cc, _, cst = oracle.CpuScene(bench_scene).render(cam, p, rng=oracle.RNG_PCG)
    assert (gc == cc).mean() >= 0.999, (gc == cc).mean()
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.99, (rel < 1e-3).mean()
    assert abs(int(st.samples) - int(cst.samples)) <= 0.001 * cst.samples


# ---- C3 / C4 --------------------------------------------------------------

@pytest.fixture(scope="module")
def frames_1080p(ctx, bench_scene):
    """The bench view at 1080p: 4- and 16-spp frames (frame seed 7), the
    G-buffer, and a 64-spp FP64 accumulation (frame seed 1000)."""
    ctx.upload_scene(bench_scene)
    W, H = 1920, 1080
    n = W * H
    cam = bench_scene.camera(W / H)
    d = {k: ctx.alloc(12 * n) for k in ("pix", "var", "n", "p", "a", "s", "o")}
    d["cnt"] = ctx.alloc(4 * n)
    d["sum"] = ctx.alloc(24 * n)
    d["sum"].zero()
    ctx.render_accum64_d(cam, pa.params(W, H, 64, 8, frame_seed=1000), 0, d["sum"])
    ctx.gbuffer_d(cam, W, H, d["n"], d["p"], d["a"])
    ctx.sync()
    host = {k: d[k].download(np.float32, 3 * n).reshape(n, 3) for k in ("n", "p", "a")}
    frames = {}
    for spp in (4, 16):
        ctx.render_d(cam, pa.params(W, H, spp, 8, tile=16, frame_seed=7), d["pix"], d["cnt"], d["var"])
        ctx.sync()
        frames[spp] = d["pix"].download(np.float32, 3 * n).reshape(n, 3)
    yield W, H, cam, d, host, frames
    for b in d.values():
        b.free()


@pytest.mark.parametrize("spp", [4, 16])
def test_c3_atrous_modes_on_the_1080p_frame(ctx, frames_1080p, spp):
    W, H, _, d, g, frames = frames_1080p
    n = W * H
    d["pix"].upload(frames[spp])
    ctx.atrous_fixed_d(d["pix"], d["n"], d["p"], d["a"], W, H, d["s"], d["o"])
    ctx.sync()
    got = d["o"].download(np.float32, 3 * n).reshape(n, 3)
    want = oracle.atrous_fixed(frames[spp], g["n"], g["p"], g["a"], W, H, 5, pa.ATROUS_FIXED_SIGMAS)
    _same_or_both_nan(got, want)
    ctx.atrous_d(d["pix"], d["n"], d["p"], d["a"], W, H, d["s"], d["o"], tile=(16, 16))
    ctx.sync()
    got = d["o"].download(np.float32, 3 * n).reshape(n, 3)
    want = oracle.atrous(frames[spp], g["n"], g["p"], g["a"], W, H, (16, 16), 5)
    _same_or_both_nan(got, want)


@pytest.mark.parametrize("spp", [4, 16])
def test_c3_relmse_of_the_1080p_frame(ctx, frames_1080p, spp):
    W, H, _, d, _, frames = frames_1080p
    n = W * H
    d["pix"].upload(frames[spp])
    got = ctx.relmse_d(d["pix"], d["sum"], 64.0, n)
    ref = d["sum"].download(np.float64, 3 * n).reshape(n, 3) / 64.0
    want = float(((frames[spp].astype(np.float64) - ref) ** 2 / (ref * ref + 1e-2)).mean())
    assert abs(got - want) <= 1e-12 * want, (got, want)


def test_c4_learned_filter_on_the_1080p_frame(ctx, frames_1080p):
    W, H, _, d, _, frames = frames_1080p
    n = W * H
    d["pix"].upload(frames[4])
    net = pa.CNN(ctx, W, H, 1)
    net.load()
    net.denoise_gbuffer_d(d["pix"], d["n"], d["p"], d["a"], W, H, d["o"])
    ctx.sync()
    out = d["o"].download(np.float32, 3 * n).reshape(H, W, 3)
    net.close()
    assert np.isfinite(out).all()
    # softmax weights: each output is a convex combination of the clamped 5x5
    # neighbourhood of the input radiance, per channel
    rad = np.pad(frames[4].reshape(H, W, 3), ((2, 2), (2, 2), (0, 0)), mode="edge")
    win = np.lib.stride_tricks.sliding_window_view(rad, (5, 5), axis=(0, 1))
    lo, hi = win.min(axis=(-2, -1)), win.max(axis=(-2, -1))
    tol = 1e-5 * np.maximum(np.abs(lo), np.abs(hi)) + 1e-6
    assert ((out >= lo - tol) & (out <= hi + tol)).all()


def test_c4_learned_filter_against_the_oracle_480x270(ctx, bench_scene):
    ctx.upload_scene(bench_scene)
    W, H = 480, 270
    n = W * H
    cam = bench_scene.camera(W / H)
    d = {k: ctx.alloc(12 * n) for k in ("pix", "var", "n", "p", "a", "o")}
    d["cnt"], d["in"] = ctx.alloc(4 * n), ctx.alloc(40 * n)
    ctx.render_d(cam, pa.params(W, H, 4, 8, frame_seed=7), d["pix"], d["cnt"], d["var"])
    ctx.gbuffer_d(cam, W, H, d["n"], d["p"], d["a"])
    ctx.cnn_inputs_d(d["pix"], d["n"], d["p"], d["a"], W, H, d["in"])
    net = pa.CNN(ctx, W, H, 1)
    net.load()
    net.denoise_d(d["in"], W, H, d["o"])
    ctx.sync()
    got = d["o"].download(np.float32, 3 * n).reshape(n, 3)
    x = d["in"].download(np.float32, 10 * n).reshape(10, H, W)
    p = np.load(pa.weights_path(), allow_pickle=False).astype(np.float64)
    want = oracle.cnn_forward(p, x, H, W)
    net.close()
    for b in d.values():
        b.free()
    err = np.abs(got - want) / (np.abs(want) + 1e-2)
    assert np.median(err) < 5e-3 and err.max() < 0.1, (np.median(err), err.max())


# ---- C5 -------------------------------------------------------------------

def test_c5_64_views_equal_single_views(ctx, bench_scene):
    from bench import look_at_box
    ctx.upload_scene(bench_scene)
    rng = np.random.default_rng(3)
    sizes = rng.choice([128, 256, 512, 1024], 64)
    spps = rng.integers(4, 65, 64)
    views, bufs = [], []
    for i in range(64):
        w = h = int(sizes[i])
        b = (ctx.alloc(12 * w * h), ctx.alloc(4 * w * h), ctx.alloc(12 * w * h))
        v = pa.View()
        v.camera = look_at_box(rng)
        v.params = pa.params(w, h, int(spps[i]), 8, tile=16, frame_seed=11, view_id=i)
        v.d_pixels, v.d_sample_count, v.d_variance = b[0].ptr.value, b[1].ptr.value, b[2].ptr.value
        views.append(v)
        bufs.append(b)
    st = ctx.render_views_d(views)
    assert sum(s.samples for s in st) > 0
    batched = [(b[0].download(np.float32, 3 * int(sizes[i]) ** 2), b[1].download(np.uint32, int(sizes[i]) ** 2))
               for i, b in enumerate(bufs)]
    big = 1024 * 1024
    sp, sc, sv = ctx.alloc(12 * big), ctx.alloc(4 * big), ctx.alloc(12 * big)
    for i, v in enumerate(views):
        n = int(sizes[i]) ** 2
        ctx.render_d(v.camera, v.params, sp, sc, sv)
        assert np.array_equal(sc.download(np.uint32, n), batched[i][1]), i
        assert np.array_equal(sp.download(np.float32, 3 * n), batched[i][0]), i
    # the render + denoise launch gives each view what the separate calls give
    outs = [ctx.alloc(12 * int(sizes[i]) ** 2) for i in range(64)]
    net = pa.CNN(ctx, 1024, 1024, 1)
    net.load()
    st = ctx.render_views_denoise_d(views, net, outs)
    assert all(s.denoise_fallback == 0 for s in st)
    gn, gp, ga, go = (ctx.alloc(12 * big) for _ in range(4))
    for i, v in enumerate(views):
        w = h = int(sizes[i])
        ctx.render_d(v.camera, v.params, sp, sc, sv)
        ctx.gbuffer_d(v.camera, w, h, gn, gp, ga)
        net.denoise_gbuffer_d(sp, gn, gp, ga, w, h, go)
        ctx.sync()
        assert np.array_equal(go.download(np.float32, 3 * w * h), outs[i].download(np.float32, 3 * w * h)), i
    net.close()
    for b in bufs:
        for x in b:
            x.free()
    for x in outs + [sp, sc, sv, gn, gp, ga, go]:
        x.free()
