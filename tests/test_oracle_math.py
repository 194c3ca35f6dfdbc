"""The oracle against what mathematics fixes: published known-answer vectors,
closed forms, invariants and brute force.  These pin the parts of the oracle
that have no reference counterpart (PCG32, GGX) and re-pin the rest."""
import ctypes as C

import numpy as np
import pytest

import oracle
import pico_amd as pa

lib = oracle.cpu_lib


def test_pcg32_published_sequence():
    # pcg32-demo (O'Neill, pcg-random.org): pcg32_srandom_r(&rng, 42u, 54u)
    out = np.empty(6, np.uint32)
    lib().cpuref_pcg32(42, 54, 6, oracle._p(out))
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


@pytest.mark.parametrize("delta", [0, 1, 2, 7, 100, 1000])
def test_pcg32_advance_equals_stepping(delta):
    a = np.empty(delta + 5, np.uint32)
    lib().cpuref_pcg32(123, 7, delta + 5, oracle._p(a))
    b = np.empty(5, np.uint32)
    lib().cpuref_pcg32_advance(123, 7, delta, 5, oracle._p(b))
    assert np.array_equal(a[delta:], b)


def test_radical_inverse_known_values():
    bits = np.array([0, 1, 2, 3, 4, 0x80000000, 0xffffffff], np.uint32)
    out = np.empty((7, 2), np.float32)
    lib().cpuref_hammersley(7, oracle._p(bits), oracle._p(out))
    assert list(out[:5, 1]) == [0.0, 0.5, 0.25, 0.75, 0.125]
    assert out[5, 1] == np.float32(2.0 ** -32)
    # first coordinate is i / 2^32 (float(UINT32_MAX) rounds to 2^32)
    assert out[5, 0] == 0.5 and out[6, 0] == 1.0


def _pdf(kind, h, r):
    n = h.shape[0]
    out = np.empty(n, np.float32)
    wo = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    hh = np.ascontiguousarray(h, np.float32)
    rr = np.full(n, r, np.float32)
    lib().cpuref_dist_pdf(kind, n, oracle._p(wo), oracle._p(hh), oracle._p(rr), oracle._p(out))
    return out.astype(np.float64)


def _sample(kind, n, r, seed, v=(0.0, 0.0, 1.0)):
    rng = np.random.default_rng(seed)
    xi = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    vv = np.tile(np.array(v, np.float32), (n, 1))
    h = np.empty((n, 3), np.float32)
    rr = np.full(n, r, np.float32)
    lib().cpuref_dist_sample(kind, n, oracle._p(xi), oracle._p(vv), oracle._p(rr), oracle._p(h))
    return h.astype(np.float64)


def _hemisphere_grid(nt=2000, npz=64):
    # midpoint rule in (cos theta, phi): d omega = d cos d phi
    ct = (np.arange(nt) + 0.5) / nt
    ph = (np.arange(npz) + 0.5) / npz * 2 * np.pi
    CT, PH = np.meshgrid(ct, ph, indexing="ij")
    st = np.sqrt(1 - CT ** 2)
    h = np.stack([st * np.cos(PH), st * np.sin(PH), CT], axis=-1).reshape(-1, 3)
    return h, (1.0 / nt) * (2 * np.pi / npz)


@pytest.mark.parametrize("r", [0.3, 0.5, 0.7, 0.9])
def test_ggx_pdf_integrates_to_one(r):
    # D(h) cos(theta_h) integrates to 1 over the hemisphere (Walter et al. 2007)
    h, dw = _hemisphere_grid()
    assert abs(_pdf(oracle.DIST_GGX, h, r).sum() * dw - 1.0) < 2e-3


def test_cosine_pdf_integrates_to_one_and_mean():
    h, dw = _hemisphere_grid()
    assert abs(_pdf(oracle.DIST_COS, h, 0.5).sum() * dw - 1.0) < 1e-3
    s = _sample(oracle.DIST_COS, 200000, 0.5, 1)
    assert abs(s[:, 2].mean() - 2.0 / 3.0) < 3e-3  # E[cos] = 2/3
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-5)


@pytest.mark.parametrize("kind,r", [(oracle.DIST_GGX, 0.4), (oracle.DIST_GGX, 0.8), (oracle.DIST_COS, 0.5)])
def test_sampler_matches_pdf_chi_square(kind, r):
    n = 200000
    s = _sample(kind, n, r, 7)
    edges = np.linspace(0, 1, 41)
    counts, _ = np.histogram(s[:, 2], bins=edges)
    # expected mass per cos-theta bin: integral of pdf over the bin (isotropic)
    exp = []
    for a, b in zip(edges[:-1], edges[1:]):
        ct = np.linspace(a, b, 201)[:-1] + (b - a) / 400
        st = np.sqrt(1 - ct ** 2)
        h = np.stack([st, np.zeros_like(ct), ct], axis=1)
        exp.append(_pdf(kind, h, r).mean() * (b - a) * 2 * np.pi)
    exp = np.array(exp) * n
    m = exp > 20
    chi2 = (((counts - exp) ** 2) / exp)[m].sum()
    dof = m.sum() - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


def test_ggx_small_roughness_is_mirror():
    s = _sample(oracle.DIST_GGX, 2000, 0.02, 3)
    assert s[:, 2].min() > 0.999


def test_samples_follow_view_hemisphere():
    for kind in (oracle.DIST_GGX, oracle.DIST_BECKMANN):
        s = _sample(kind, 5000, 0.5, 4, v=(0.3, 0.1, -0.9))
        assert (s[:, 2] < 0).all()


def test_beckmann_all_density_at_normal():
    # D(h = n) = 1 / (pi alpha^2) before the clamp to [0, 1]; alpha is PBRT's
    # roughness-to-alpha polynomial of log(max(R, 1e-3))
    for r in [0.5, 0.8, 1.0]:
        x = np.log(np.float32(r))
        alpha = 1.62142 + 0.819955 * x + 0.1734 * x * x + 0.0171201 * x ** 3 + 0.000640711 * x ** 4
        want = min(1.0, 1.0 / (np.pi * alpha * alpha))
        got = _pdf(oracle.DIST_BECKMANN, np.array([[0.0, 0.0, 1.0]]), r)[0]
        assert abs(got - want) < 1e-5 * max(1.0, want)


def test_alias_table_reproduces_weights():
    rng = np.random.default_rng(2)
    for w in [np.ones(7, np.float32), np.array([0, 0, 1, 0], np.float32),
              rng.uniform(0.1, 5, 50).astype(np.float32)]:
        n = len(w)
        bw, ba = np.empty(n, np.float32), np.empty(n, np.uint32)
        lib().cpuref_alias_build(n, oracle._p(w), oracle._p(bw), oracle._p(ba))
        # probability of outcome j: sum over buckets of (1/n)(w_b [b==j] + (1-w_b)[alias_b==j])
        prob = np.zeros(n)
        for b in range(n):
            prob[b] += bw[b] / n
            prob[ba[b]] += (1 - bw[b]) / n
        np.testing.assert_allclose(prob, w / w.sum(), atol=1e-5)


def test_bvh_matches_brute_force():
    s = pa.Scene.instanced(3, 4, 40, 100, 800, 10.0)
    cs = oracle.CpuScene(s)
    rng = np.random.default_rng(0)
    n = 3000
    o = rng.uniform(-2, 12, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t, inst, prim, *_ = cs.closest_hit(o, d)
    bt, binst, bprim = cs.closest_hit(o, d, brute=True)
    assert (inst != 0xffffffff).mean() > 0.2
    assert np.array_equal(inst, binst) and np.array_equal(prim, bprim)
    np.testing.assert_array_equal(t, bt)


def test_camera_centre_and_corner_rays():
    cam = pa.Camera((0.5, -1.0, 2.0), (0.0, 0.0, 2.0), (0.0, 1.0, 0.0), 2.0, 100.0)
    px = np.array([64, 0], np.uint32)
    py = np.array([32, 0], np.uint32)
    xi = np.zeros((2, 2), np.float32)
    d = np.empty((2, 3), np.float32)
    lib().cpuref_camera_rays(cam, 128, 64, 2, oracle._p(px), oracle._p(py), oracle._p(xi), oracle._p(d))
    np.testing.assert_allclose(d[0], [0, 0, 1], atol=1e-7)
    # pixel (0,0): (-0.5*aspect)*right + (-0.5)*up + 1*dir, right = cross(dir, up) = (-1, 0, 0)
    want = np.array([1.0, -0.5, 2.0]) / np.linalg.norm([1.0, -0.5, 2.0])
    np.testing.assert_allclose(d[1], want, atol=1e-6)


def _single_light_scene():
    """One emissive quad filling the view, black sky: every camera ray that
    hits the light returns exactly L_e (emission added at depth 0, the path then
    leaves into the black sky)."""
    return pa.Scene.cornell(1, pa.MICROFACET_GGX)


def test_depth_zero_returns_black_for_hits_and_sky_for_misses():
    s = pa.Scene.instanced(5, 2, 20, 100, 300, 8.0)
    cam = s.camera(1.0)
    cs = oracle.CpuScene(s)
    p = pa.params(32, 32, 2, 0, frame_seed=1)
    out, _ = cs.render_samples(cam, p)
    inst, _ = cs.primary_ids(cam, 32, 32, (0.0, 0.0))
    sky = np.array(s.desc.sky_rgb[:])
    miss = out.reshape(-1, 3)[np.tile(inst, 2) == 0xffffffff]
    hit = out.reshape(-1, 3)[np.tile(inst, 2) != 0xffffffff]
    # camera jitter can change a hit to a miss at silhouettes: compare sets, not pixels
    assert np.isin(miss, [0.0, sky[0]]).all()
    assert ((hit == 0) | (hit == sky[0])).all()


def test_light_seen_directly_returns_emission():
    s = pa.Scene.cornell(1, pa.MICROFACET_GGX)
    cs = oracle.CpuScene(s)
    # camera under the light looking straight up at it
    cam = pa.Camera((0.0, 0.5, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 1.0, 100.0)
    p = pa.params(16, 16, 3, 1, frame_seed=2)
    out, _ = cs.render_samples(cam, p)
    inst, _ = cs.primary_ids(cam, 16, 16, (0.0, 0.0))
    light = s.desc.n_instances - 1
    sel = np.tile(inst == light, 3)
    # depth 1: emission at the light, NEE is skipped at lights, the extension
    # ray's hit at depth 1 is not shaded
    np.testing.assert_array_equal(out.reshape(-1, 3)[sel], 10.0)


def test_diffuse_plane_under_constant_sky():
    """A camera ray hits a diffuse surface facing an open constant sky L: the
    reference's rule throughput *= pdf * energy gives E[radiance] = rho L E[cos]/pi
    = 2 rho L / (3 pi) for one bounce (cosine-sampled, no lights, no occluder)."""
    s = pa.Scene.instanced(9, 2, 1, 50, 60, 1.0)
    # use the restatement's render on a floor quad: build a tiny description by hand
    pos = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    vofs = np.array([0, 4], np.uint32)
    tofs = np.array([0, 2], np.uint32)
    xf = np.eye(4, dtype=np.float32).reshape(-1)
    im = np.zeros(1, np.uint32)
    rho = 0.6
    mats = (pa.Material * 1)(pa.Material((rho, rho, rho), (0, 0, 0), 0.5, (0, 0, 0)))
    bsdf = np.array([pa.BSDF_DIFFUSE], np.uint32)
    d = pa.SceneDesc()
    d.n_meshes, d.mesh_vertex_offset, d.mesh_triangle_offset = 1, vofs.ctypes.data, tofs.ctypes.data
    d.positions, d.normals, d.indices = pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data
    d.n_instances, d.transforms, d.instance_mesh = 1, xf.ctypes.data, im.ctypes.data
    d.instance_material, d.instance_bsdf = im.ctypes.data, bsdf.ctypes.data
    d.n_materials, d.materials = 1, C.cast(mats, C.c_void_p).value
    d.n_lights, d.light_instances = 0, None
    L = 2.0
    d.sky_rgb[:] = [L, L, L]

    class Holder:
        desc = d
    s = Holder()
    cs = oracle.CpuScene(s)
    cam = pa.Camera((0.0, 2.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), 1.0, 100.0)
    p = pa.params(32, 32, 64, 1, frame_seed=4)
    out, _ = cs.render_samples(cam, p)
    want = rho * L * 2.0 / (3.0 * np.pi)
    se = out[..., 0].std() / np.sqrt(out[..., 0].size)
    assert abs(out[..., 0].mean() - want) < 4 * se
