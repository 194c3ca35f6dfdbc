"""The oracle and the device path against golden vectors of Pico's own
hot-path functions (tests/golden/ref_vectors.npz, written by
tests/golden/make_golden.py from the reference subset compiled from the Pico
source tree).  These run wherever the compiled subset is absent."""
import os

import numpy as np
import pytest

import oracle
import pico_amd as pa
from sky_scenes import shapes_sun_sky, sun_cube_sky, sun_hdri_sky

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"), allow_pickle=False)
P = oracle._p
SCENES = ["cornell_beckmann", "instanced_small", "cornell_glass", "shapes", "sun_cube_sky", "sun_hdri_sky",
          "shapes_sun_sky"]


def _scene(name):
    if name == "sun_cube_sky":
        return sun_cube_sky()
    if name == "sun_hdri_sky":
        return sun_hdri_sky()
    if name == "shapes_sun_sky":
        return shapes_sun_sky()
    if name == "cornell_beckmann":
        return pa.Scene.cornell(1, pa.MICROFACET_BECKMANN_ALL)
    if name == "cornell_glass":
        return pa.Scene.cornell_glass(1, pa.MICROFACET_BECKMANN_ALL)
    if name == "shapes":
        return pa.Scene.shapes(1, pa.MICROFACET_BECKMANN_ALL)
    return pa.Scene.instanced(7, 4, 48, 200, 2000, 12.0, pa.MICROFACET_BECKMANN_ALL)


@pytest.fixture(scope="module")
def scenes():
    out = {}
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                             "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for name in SCENES:
        s = _scene(name)
        # the procedural generator must still produce the scene the vectors were made on
        assert mg.scene_hash(s).encode() == G[f"{name}_hash"].tobytes()
        out[name] = s
    return out


def test_generators():
    for seed in [0, 1, 12345, 0xdeadbeef]:
        a = np.empty(256, np.uint32)
        oracle.cpu_lib().cpuref_xorshift(seed, 256, P(a))
        assert np.array_equal(a, G[f"xorshift_{seed}"])
    h = np.empty((256, 2), np.float32)
    bits = np.ascontiguousarray(G["hammersley_in"])  # bound: the pointer must outlive the call
    oracle.cpu_lib().cpuref_hammersley(256, P(bits), P(h))
    assert np.array_equal(h, G["hammersley_out"])


@pytest.mark.parametrize("kind,name", [(oracle.DIST_COS, "cos"), (oracle.DIST_BECKMANN, "beckmann")])
def test_distributions(kind, name):
    xi, v, r = (np.ascontiguousarray(G[f"dist_{name}_{k}"]) for k in ("xi", "v", "r"))
    n = len(r)
    h, pdf = np.empty((n, 3), np.float32), np.empty(n, np.float32)
    oracle.cpu_lib().cpuref_dist_sample(kind, n, P(xi), P(v), P(r), P(h))
    np.testing.assert_allclose(h, G[f"dist_{name}_h"], rtol=2e-6, atol=2e-7)
    hr = np.ascontiguousarray(G[f"dist_{name}_h"])
    oracle.cpu_lib().cpuref_dist_pdf(kind, n, P(v), P(hr), P(r), P(pdf))
    np.testing.assert_allclose(pdf, G[f"dist_{name}_pdf"], rtol=1e-5, atol=1e-7)


def test_tangent_frame():
    v, nn = np.ascontiguousarray(G["frame_v"]), np.ascontiguousarray(G["frame_n"])
    out = np.empty((len(v), 9), np.float32)
    oracle.cpu_lib().cpuref_tangent_frame(len(v), P(v), P(nn), P(out))
    np.testing.assert_allclose(out, G["frame_out"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("k", [1, 2, 5, 64])
def test_alias_table(k):
    w = np.ascontiguousarray(G[f"alias_{k}_w"])
    bw, ba = np.empty(k, np.float32), np.empty(k, np.uint32)
    oracle.cpu_lib().cpuref_alias_build(k, P(w), P(bw), P(ba))
    s = np.empty(512, np.uint32)
    oracle.cpu_lib().cpuref_alias_sample(k, P(bw), P(ba), 99, 512, P(s))
    assert np.array_equal(s, G[f"alias_{k}_samples"])


@pytest.mark.parametrize("bsdf", [pa.BSDF_DIFFUSE, pa.BSDF_SPECULAR, pa.BSDF_DIELECTRIC, pa.BSDF_LIGHT, pa.BSDF_DELTA])
def test_bsdf_sample(bsdf):
    v, nrm = np.ascontiguousarray(G[f"bsdf_{bsdf}_v"]), np.ascontiguousarray(G[f"bsdf_{bsdf}_n"])
    n = len(v)
    m = pa.Material((0.4, 0.5, 0.6), (0.04, 0.04, 0.04), 0.35, (2.0, 2.0, 2.0))
    l, p, e = np.empty((n, 3), np.float32), np.empty(n, np.float32), np.empty((n, 3), np.float32)
    oracle.cpu_lib().cpuref_bsdf_sample(bsdf, pa.MICROFACET_BECKMANN_ALL, m, n, P(v), P(nrm), oracle.RNG_REFERENCE, 777,
                                        P(l), P(p), P(e))
    np.testing.assert_allclose(l, G[f"bsdf_{bsdf}_l"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(p, G[f"bsdf_{bsdf}_p"], rtol=1e-4, atol=1e-6)
    np.testing.assert_array_equal(e, G[f"bsdf_{bsdf}_e"])


@pytest.mark.parametrize("bsdf", [pa.BSDF_TRANSPARENT, pa.BSDF_FRESNEL])
def test_btdf_sample(bsdf):
    v, nrm = np.ascontiguousarray(G[f"btdf_{bsdf}_v"]), np.ascontiguousarray(G[f"btdf_{bsdf}_n"])
    inside = np.ascontiguousarray(G[f"btdf_{bsdf}_inside"])
    n = len(v)
    m = pa.Material((0.9, 0.95, 1.0), (0.5, 0.5, 0.5), 0.3, (0, 0, 0), 1.45)
    l, p, e = np.empty((n, 3), np.float32), np.empty(n, np.float32), np.empty((n, 3), np.float32)
    d = np.empty(n, np.uint32)
    oracle.cpu_lib().cpuref_btdf_sample(bsdf, pa.MICROFACET_BECKMANN_ALL, m, n, P(v), P(nrm), P(inside),
                                        oracle.RNG_REFERENCE, 91, P(l), P(p), P(e), P(d))
    assert np.array_equal((d > 1).astype(np.uint8), G[f"btdf_{bsdf}_after"])
    ok = ~np.isnan(G[f"btdf_{bsdf}_l"]).any(axis=1)
    assert np.array_equal(np.isnan(l).any(axis=1), ~ok)
    np.testing.assert_allclose(l[ok], G[f"btdf_{bsdf}_l"][ok], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(p[ok], G[f"btdf_{bsdf}_p"][ok], rtol=1e-4, atol=1e-6)
    np.testing.assert_array_equal(e, G[f"btdf_{bsdf}_e"])


def test_camera_rays():
    cam = pa.Camera((1.0, 2.0, -3.0), (0.2, -0.1, 1.0), (0.0, 1.0, 0.0), 1.5, 1000.0)
    px, py, xi = (np.ascontiguousarray(G[f"camera_{k}"]) for k in ("px", "py", "xi"))
    d = np.empty((len(px), 3), np.float32)
    oracle.cpu_lib().cpuref_camera_rays(cam, 320, 200, len(px), P(px), P(py), P(xi), P(d))
    assert np.array_equal(d, G["camera_dir"])


@pytest.mark.parametrize("name", SCENES)
def test_closest_hit_and_ids(scenes, name):
    cs = oracle.CpuScene(scenes[name])
    inst, prim = cs.primary_ids(scenes[name].camera(1.0), 48, 48, (0.25, 0.75))
    assert np.array_equal(inst, G[f"{name}_ids_inst"]) and np.array_equal(prim, G[f"{name}_ids_prim"])
    t, hi, hp, _, pos, nrm = cs.closest_hit(G[f"{name}_ray_o"], G[f"{name}_ray_d"])
    assert np.array_equal(hi, G[f"{name}_hit_inst"]) and np.array_equal(hp, G[f"{name}_hit_prim"])
    m = np.isfinite(G[f"{name}_hit_t"])
    np.testing.assert_allclose(t[m], G[f"{name}_hit_t"][m], rtol=1e-5, atol=5e-5)
    np.testing.assert_allclose(pos[m], G[f"{name}_hit_pos"][m], rtol=1e-5, atol=5e-5)
    np.testing.assert_allclose(nrm[m], G[f"{name}_hit_nrm"][m], rtol=1e-5, atol=5e-5)


@pytest.mark.parametrize("name", SCENES)
def test_integrator_samples(scenes, name):
    p = pa.params(16, 16, 6, 5, frame_seed=21)
    a, _ = oracle.CpuScene(scenes[name]).render_samples(scenes[name].camera(1.0), p, rng=oracle.RNG_REFERENCE)
    b = G[f"{name}_samples"]
    rel = (np.abs(a - b) / (np.abs(b) + 1e-6)).max(axis=2)
    assert (rel < 1e-4).mean() >= 0.995, (rel < 1e-4).mean()


def test_accumulation(scenes):
    s = scenes["cornell_beckmann"]
    p = pa.params(16, 16, 12, 4, frame_seed=5)
    pix, cnt, var, _ = oracle.CpuScene(s).render(s.camera(1.0), p, rng=oracle.RNG_REFERENCE, threads=1)
    assert np.array_equal(cnt, G["accum_cnt"])
    np.testing.assert_allclose(pix, G["accum_pix"], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(var, G["accum_var"], rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize("name", SCENES)
def test_gbuffer(scenes, name):
    out = oracle.CpuScene(scenes[name]).gbuffer(scenes[name].camera(1.0), 32, 24)
    for x, k in zip(out, ("n", "p", "a")):
        np.testing.assert_allclose(x, G[f"{name}_gb_{k}"], rtol=1e-5, atol=5e-4)


@pytest.mark.parametrize("name", ["cube", "hdri"])
def test_sky_lookup(name):
    import ctypes as C
    tex, d = np.ascontiguousarray(G[f"sky_{name}_tex"]), np.ascontiguousarray(G[f"sky_{name}_dir"])
    img = pa.SkyImage(tex.shape[2], tex.shape[1], tex.shape[0],
                      pa.TEXEL_RGBA32F if tex.dtype == np.float32 else pa.TEXEL_RGBA8, tex.ctypes.data)
    sky = (C.c_float * 3)(0.0, 0.0, 0.0)
    out = np.empty((len(d), 3), np.float32)
    oracle.cpu_lib().cpuref_sky_sample(C.addressof(img), C.addressof(sky), len(d), P(d), P(out))
    assert np.array_equal(out, G[f"sky_{name}_out"])


@pytest.mark.parametrize("tile", [(16, 16), (5, 7)])
def test_atrous(tile):
    args = [G[f"atrous_{k}"] for k in ("rgb", "n", "p", "a")]
    out = oracle.atrous(*args, 41, 31, tile, 5)
    np.testing.assert_allclose(out, G[f"atrous_out_{tile[0]}x{tile[1]}"], rtol=1e-6, atol=1e-7)


def test_tonemap():
    np.testing.assert_array_equal(oracle.tonemap(G["tonemap_in"], 64, 32), G["tonemap_out"])


# ---- the device path against the same vectors -------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name", SCENES)
def test_device_closest_hit_and_ids(ctx, scenes, name):
    ctx.upload_scene(scenes[name])
    t, hi, hp, _ = ctx.closest_hit(G[f"{name}_ray_o"], G[f"{name}_ray_d"])
    assert np.array_equal(hi, G[f"{name}_hit_inst"]) and np.array_equal(hp, G[f"{name}_hit_prim"])
    m = np.isfinite(G[f"{name}_hit_t"])
    np.testing.assert_allclose(t[m], G[f"{name}_hit_t"][m], rtol=1e-5, atol=5e-5)
    inst, prim = ctx.primary_ids(scenes[name].camera(1.0), 48, 48, (0.25, 0.75))
    same = (inst == G[f"{name}_ids_inst"]) & (prim == G[f"{name}_ids_prim"])
    assert same.mean() >= 0.9999 or (~same).sum() <= 1, (~same).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("name", SCENES)
def test_device_gbuffer(ctx, scenes, name):
    ctx.upload_scene(scenes[name])
    out = ctx.gbuffer(scenes[name].camera(1.0), 32, 24)
    close = [np.isclose(x, G[f"{name}_gb_{k}"], rtol=1e-5, atol=5e-4).all(axis=1) for x, k in zip(out, ("n", "p", "a"))]
    # every position matches; a normal or albedo may differ only where the
    # camera ray meets two surfaces at the same point (an edge shared by two
    # walls), a tie that the traversal order resolves
    assert close[1].all()
    other = ~(close[0] & close[2])
    assert other.sum() <= 2, np.where(other)[0]


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [(16, 16), (5, 7)])
def test_device_atrous(ctx, tile):
    args = [G[f"atrous_{k}"] for k in ("rgb", "n", "p", "a")]
    out = ctx.atrous(*args, 41, 31, tile, 5)
    np.testing.assert_allclose(out, G[f"atrous_out_{tile[0]}x{tile[1]}"], rtol=1e-4, atol=1e-6)


@pytest.mark.gpu
def test_device_tonemap(ctx):
    np.testing.assert_allclose(ctx.tonemap(G["tonemap_in"], 64, 32), G["tonemap_out"], rtol=1e-6, atol=1e-7)
