"""Device parity: the HIP path (through the C ABI) against the oracle."""
import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cornell():
    return pa.Scene.cornell(1, pa.MICROFACET_GGX)


@pytest.fixture(scope="module")
def instanced():
    return pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_GGX)


def _random_rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_closest_hit_cornell(ctx, cornell):
    ctx.upload_scene(cornell)
    o, d = _random_rays(20000, 1, -0.95, 0.95)
    t, inst, prim, uv = ctx.closest_hit(o, d)
    ct, ci, cp, *_ = oracle.CpuScene(cornell).closest_hit(o, d)
    agree = (inst == ci) & (prim == cp)
    assert agree.mean() >= 0.9999, agree.mean()
    m = agree & np.isfinite(ct)
    np.testing.assert_allclose(t[m], ct[m], rtol=1e-4, atol=1e-5)


def test_closest_hit_instanced_full_scene(ctx, instanced):
    ctx.upload_scene(instanced)
    o, d = _random_rays(50000, 2, 0.0, 100.0)
    t, inst, prim, uv = ctx.closest_hit(o, d)
    ct, ci, cp, *_ = oracle.CpuScene(instanced).closest_hit(o, d)
    assert (ci != 0xffffffff).mean() > 0.3
    agree = (inst == ci) & (prim == cp)
    assert agree.mean() >= 0.9999, agree.mean()
    m = agree & np.isfinite(ct)
    # the reference's distance is |M p(u,v) - o| of the interpolated hit point,
    # the device's the ray parameter: they differ where a ray grazes a triangle
    close = np.isclose(t[m], ct[m], rtol=1e-4, atol=1e-4)
    assert close.mean() >= 0.9999


@pytest.mark.parametrize("xi", [(0.37, 0.61), (0.9, 0.1)])
def test_primary_ids_jittered(ctx, cornell, instanced, xi):
    for scene, w, h in [(cornell, 256, 256), (instanced, 480, 270)]:
        ctx.upload_scene(scene)
        cam = scene.camera(w / h)
        gi, gp = ctx.primary_ids(cam, w, h, xi)
        ci, cp = oracle.CpuScene(scene).primary_ids(cam, w, h, xi)
        assert ((gi == ci) & (gp == cp)).mean() >= 0.9999


def test_pixel_corner_rays_differ_only_at_exact_ties(ctx, cornell):
    """xi = (0,0) (the G-buffer rays) puts rays of the Cornell box exactly on its
    wall-to-wall corners and on the shared diagonals of its quads: two surfaces
    at the same distance, and traversal order picks one (first found in the
    reference).  Every disagreement must be such a tie: same hit distance."""
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    w = h = 256
    px = np.tile(np.arange(w, dtype=np.uint32), h)
    py = np.repeat(np.arange(h, dtype=np.uint32), w)
    xi = np.zeros((w * h, 2), np.float32)
    d = np.empty((w * h, 3), np.float32)
    oracle.cpu_lib().cpuref_camera_rays(cam, w, h, w * h, oracle._p(px), oracle._p(py), oracle._p(xi), oracle._p(d))
    o = np.tile(np.array(cam.position[:], np.float32), (w * h, 1))
    t, gi, gp, _ = ctx.closest_hit(o, d)
    ct, ci, cp, *_ = oracle.CpuScene(cornell).closest_hit(o, d)
    same = (gi == ci) & (gp == cp)
    assert same.mean() >= 0.99
    np.testing.assert_allclose(t[~same], ct[~same], rtol=1e-5)
    gi2, gp2 = ctx.primary_ids(cam, w, h, (0.0, 0.0))
    assert np.array_equal(gi2, gi) and np.array_equal(gp2, gp)


def test_render_samples_match_oracle_pathwise(ctx, cornell):
    ctx.upload_scene(cornell)
    cam = cornell.camera(1.0)
    p = pa.params(64, 64, 4, 5, frame_seed=11)
    g = ctx.render_samples(cam, p)
    c, _ = oracle.CpuScene(cornell).render_samples(cam, p, rng=oracle.RNG_PCG)
    rel = (np.abs(g - c) / (np.abs(c) + 1e-3)).max(axis=2)
    # identical PCG streams: paths agree unless a rounding difference flips a
    # discrete decision (lobe, roulette, hit/miss)
    assert (rel < 1e-3).mean() >= 0.99, (rel < 1e-3).mean()


@pytest.mark.parametrize("levels", [5, 7])
def test_atrous_matches_oracle(ctx, levels):
    # 5 levels: dilations 1..16 (the LDS-staged kernel); 7 levels adds
    # dilations 32 and 64 (the direct kernel)
    rng = np.random.default_rng(5)
    w, h = 100, 75
    rgb = rng.gamma(2.0, 0.5, size=(w * h, 3)).astype(np.float32)
    n = rng.normal(size=(w * h, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pos = rng.normal(size=(w * h, 3)).astype(np.float32) * 3
    alb = rng.uniform(0, 1, size=(w * h, 3)).astype(np.float32)
    for tile in [(16, 16), (64, 64), (7, 5)]:
        g = ctx.atrous(rgb, n, pos, alb, w, h, tile, levels)
        c = oracle.atrous(rgb, n, pos, alb, w, h, tile, levels)
        np.testing.assert_allclose(g, c, rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("sigmas", [(0.5, 0.05, 0.1, 0.0), (0.3, 0.1, 0.2, 0.8), (0.0, 0.0, 0.0, 0.0)])
def test_atrous_fixed_matches_oracle(ctx, sigmas):
    rng = np.random.default_rng(6)
    w, h = 100, 75
    rgb = rng.gamma(2.0, 0.5, size=(w * h, 3)).astype(np.float32)
    n = rng.normal(size=(w * h, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    pos = rng.normal(size=(w * h, 3)).astype(np.float32) * 3
    alb = rng.uniform(0, 1, size=(w * h, 3)).astype(np.float32)
    d = [ctx.alloc(12 * w * h) for _ in range(6)]
    for arr, src in zip(d, (rgb, n, pos, alb)):
        arr.upload(src)
    ctx.atrous_fixed_d(*d[:4], w, h, d[4], d[5], levels=5, sigmas=sigmas)
    g = d[5].download(np.float32, 3 * w * h).reshape(-1, 3)
    c = oracle.atrous_fixed(rgb, n, pos, alb, w, h, 5, sigmas)
    np.testing.assert_allclose(g, c, rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("mf", [pa.MICROFACET_BECKMANN_ALL, pa.MICROFACET_GGX])
def test_transmissive_bsdfs_pathwise(ctx, mf):
    """Paths through a Fresnel glass block and a Transparent box (the IOR
    stack, offsets below the surface inside geometry, no NEE at BTDFs)
    against the oracle with identical random streams."""
    scene = pa.Scene.cornell_glass(1, mf)
    ctx.upload_scene(scene)
    cam = scene.camera(1.0)
    p = pa.params(64, 64, 4, 8, frame_seed=13)
    g = ctx.render_samples(cam, p)
    c, _ = oracle.CpuScene(scene).render_samples(cam, p, rng=oracle.RNG_PCG)
    nan_g, nan_c = np.isnan(g), np.isnan(c)
    assert np.array_equal(nan_g, nan_c)
    g, c = np.where(nan_g, 0, g), np.where(nan_c, 0, c)
    rel = (np.abs(g - c) / (np.abs(c) + 1e-3)).max(axis=2)
    assert (rel < 1e-3).mean() >= 0.995, (rel < 1e-3).mean()
    # the glass boxes are seen and transmit light
    inst, _ = ctx.primary_ids(cam, 64, 64)
    assert np.isin(inst, [5, 6]).mean() > 0.05


@pytest.mark.parametrize("mf", [pa.MICROFACET_BECKMANN_ALL, pa.MICROFACET_GGX])
def test_implicit_shapes_pathwise(ctx, mf):
    """The built-in sphere and cube (a Dielectric sphere, a metal cube, a
    Fresnel glass sphere, an emitting sphere and an emitting cube sampled by
    next-event estimation) against the oracle with identical random streams."""
    scene = pa.Scene.shapes(1, mf)
    ctx.upload_scene(scene)
    cam = scene.camera(1.0)
    p = pa.params(64, 64, 4, 8, frame_seed=17)
    g = ctx.render_samples(cam, p)
    c, _ = oracle.CpuScene(scene).render_samples(cam, p, rng=oracle.RNG_PCG)
    nan_g, nan_c = np.isnan(g), np.isnan(c)
    assert np.array_equal(nan_g, nan_c)
    g, c = np.where(nan_g, 0, g), np.where(nan_c, 0, c)
    rel = (np.abs(g - c) / (np.abs(c) + 1e-3)).max(axis=2)
    assert (rel < 1e-3).mean() >= 0.995, (rel < 1e-3).mean()
    inst, _ = ctx.primary_ids(cam, 64, 64)
    for k in (5, 6, 7):  # each shape is in view
        assert (inst == k).mean() > 0.01, k


def test_implicit_shapes_closest_hit(ctx):
    scene = pa.Scene.shapes(1, pa.MICROFACET_GGX)
    o, d = _random_rays(30000, 21, -0.9, 0.9)
    ct, ci, cp, *_ = oracle.CpuScene(scene).closest_hit(o, d)
    for device_build in (False, True):
        ctx.upload_scene(scene, device_build=device_build)
        t, inst, prim, uv = ctx.closest_hit(o, d)
        agree = (inst == ci) & (prim == cp)
        assert agree.mean() >= 0.9999, agree.mean()
        m = agree & np.isfinite(ct)
        np.testing.assert_allclose(t[m], ct[m], rtol=1e-4, atol=1e-5)
        assert np.isin(inst, [5, 6, 7, 9, 10]).mean() > 0.1
