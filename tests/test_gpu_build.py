"""Device scene build (every mesh BVH binned-SAH on the GPU, TLAS as a linear
BVH): the same binary and 4-wide trees as the host builder (node counts,
stack bound, depth), the same closest hits and images, the oracle's image,
and the error and edge cases of the description."""
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
def instanced():
    return pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_GGX)


def _random_rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def _same_hits(a, b, min_frac=0.9999):
    same = (a[1] == b[1]) & (a[2] == b[2])
    assert same.mean() >= min_frac, same.mean()
    assert np.array_equal(a[0][same], b[0][same])
    assert np.array_equal(a[3][same], b[3][same])


def _both_builds(ctx, scene, rays):
    sh = ctx.upload_scene(scene)
    host = ctx.closest_hit(*rays)
    info_h = ctx.scene_info()
    sd = ctx.upload_scene(scene, device_build=True)
    dev = ctx.closest_hit(*rays)
    info_d = ctx.scene_info()
    print(f"[build] host {sh.bvh_ms:.1f} ms / device {sd.bvh_ms:.2f} ms; binary nodes {sh.binary_nodes} / "
          f"{sd.binary_nodes}, wide {sh.blas_nodes4} / {sd.blas_nodes4}, bound {sh.blas_bound} / {sd.blas_bound}, "
          f"levels {sh.levels} / {sd.levels}", flush=True)
    return sh, sd, host, dev, info_h, info_d


@pytest.mark.parametrize("name", ["cornell", "instanced"])
def test_same_tree_and_hits_as_host_build(ctx, cornell, instanced, name):
    scene = cornell if name == "cornell" else instanced
    rays = _random_rays(50000, 3, -0.9, 0.9) if name == "cornell" else _random_rays(50000, 4, 0.0, 100.0)
    sh, sd, host, dev, info_h, info_d = _both_builds(ctx, scene, rays)
    # the same split decisions give the same binary tree and the same greedy collapse
    assert sd.binary_nodes == sh.binary_nodes
    assert sd.blas_nodes4 == sh.blas_nodes4
    assert sd.blas_bound == sh.blas_bound
    assert sd.levels == sh.levels
    assert sd.stack_bound <= 64
    assert info_d["bytes_tris"] == info_h["bytes_tris"]
    assert (host[1] != 0xffffffff).mean() > 0.3
    _same_hits(host, dev)


def test_same_image_as_host_build(ctx, instanced):
    cam = instanced.camera(16 / 9)
    p = pa.params(192, 108, 2, 6, frame_seed=4)
    ctx.upload_scene(instanced)
    img0, cnt0, _, _ = ctx.render(cam, p)
    ctx.upload_scene(instanced, device_build=True)
    img1, cnt1, _, _ = ctx.render(cam, p)
    assert np.array_equal(cnt0, cnt1)
    eq = np.all(img0 == img1, axis=1)
    assert eq.mean() >= 0.999, eq.mean()


def test_render_matches_oracle(ctx, cornell):
    ctx.upload_scene(cornell, device_build=True)
    w, h = 48, 48
    cam = cornell.camera(1.0)
    p = pa.params(w, h, 8, 5, frame_seed=6)
    gp, gc, _, _ = ctx.render(cam, p)
    cp, cc, _, _ = oracle.CpuScene(cornell).render(cam, p, rng=oracle.RNG_PCG)
    assert np.array_equal(gc, cc)
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.98, (rel < 1e-3).mean()


def test_large_mesh(ctx):
    """One mesh of about two million triangles (and its lights)."""
    scene = pa.Scene.instanced(5, 1, 4, 2000000, 2000000, 100.0, pa.MICROFACET_GGX)
    rays = _random_rays(50000, 5, 0.0, 100.0)
    sh, sd, host, dev, _, _ = _both_builds(ctx, scene, rays)
    assert sd.binary_nodes == sh.binary_nodes and sd.blas_nodes4 == sh.blas_nodes4
    assert sd.blas_bound == sh.blas_bound
    _same_hits(host, dev)


class _Desc:
    """A hand-made description; keeps its arrays alive."""

    def __init__(self, meshes, instances, lights=()):
        pos, nrm, idx, vofs, tofs = [], [], [], [0], [0]
        for p, t in meshes:
            p = np.asarray(p, np.float32).reshape(-1, 3)
            t = np.asarray(t, np.uint32).reshape(-1, 3)
            pos.append(p)
            n = np.zeros_like(p)
            n[:, 1] = 1.0
            nrm.append(n)
            idx.append(t)
            vofs.append(vofs[-1] + len(p))
            tofs.append(tofs[-1] + len(t))
        self.pos = np.ascontiguousarray(np.concatenate(pos) if pos else np.zeros((0, 3)), np.float32)
        self.nrm = np.ascontiguousarray(np.concatenate(nrm), np.float32)
        self.idx = np.ascontiguousarray(np.concatenate(idx), np.uint32)
        self.vofs, self.tofs = np.array(vofs, np.uint32), np.array(tofs, np.uint32)
        self.xf = np.ascontiguousarray([np.asarray(m, np.float32).T.reshape(-1) for m, _ in instances], np.float32)
        self.im = np.array([k for _, k in instances], np.uint32)
        self.imat = np.zeros(len(instances), np.uint32)
        self.bsdf = np.full(len(instances), pa.BSDF_DIFFUSE, np.uint32)
        self.mats = (pa.Material * 1)(pa.Material((0.5, 0.5, 0.5), (0, 0, 0), 0.5, (0, 0, 0)))
        d = pa.SceneDesc()
        d.n_meshes, d.mesh_vertex_offset, d.mesh_triangle_offset = len(meshes), self.vofs.ctypes.data, self.tofs.ctypes.data
        d.positions, d.normals, d.indices = self.pos.ctypes.data, self.nrm.ctypes.data, self.idx.ctypes.data
        d.n_instances, d.transforms, d.instance_mesh = len(instances), self.xf.ctypes.data, self.im.ctypes.data
        d.instance_material, d.instance_bsdf = self.imat.ctypes.data, self.bsdf.ctypes.data
        d.n_materials, d.materials = 1, C.cast(self.mats, C.c_void_p).value
        d.n_lights, d.light_instances = 0, None
        d.sky_rgb[:] = [1.0, 1.0, 1.0]
        self.desc = d


def _quad(y):
    return [[-1, y, -1], [1, y, -1], [1, y, 1], [-1, y, 1]], [[0, 1, 2], [0, 2, 3]]


def _translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def test_empty_mesh_and_single_leaf_meshes(ctx):
    """An empty mesh, a one-triangle mesh and a two-triangle mesh (one leaf),
    instanced side by side: both builders agree, hits match brute force."""
    one = ([[-1, 0, -1], [1, 0, -1], [0, 0, 1]], [[0, 1, 2]])
    desc = _Desc([([], []), one, _quad(0.0)],
                 [(_translate(0, 0, 0), 0), (_translate(0, 0, 0), 1), (_translate(3, 0, 0), 2), (_translate(-3, 1, 0), 1)])
    o = np.tile(np.array([[0.0, 5.0, 0.0]], np.float32), (3000, 1))
    rng = np.random.default_rng(9)
    tgt = np.stack([rng.uniform(-5, 5, 3000), np.zeros(3000), rng.uniform(-2, 2, 3000)], 1)
    d = (tgt - o).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    sh, sd, host, dev, _, _ = _both_builds(ctx, desc, (o, d))
    assert (sd.binary_nodes, sd.blas_nodes4, sd.blas_bound) == (sh.binary_nodes, sh.blas_nodes4, sh.blas_bound)
    _same_hits(host, dev, 1.0)
    bt, bi, bp, *_ = oracle.CpuScene(desc).closest_hit(o, d, brute=True)
    assert np.array_equal(dev[1], bi) and np.array_equal(dev[2], bp)


@pytest.mark.parametrize("copies", [40, 10000])
def test_degenerate_centroids(ctx, copies):
    """Copies of one triangle: every centroid coincides, both builders split
    by position in the range (10000 copies take the many-workgroup path for
    large nodes); the trees have the same shape and the hit is one of the
    copies."""
    p, t = [[-1, 0, -1], [1, 0, -1], [0, 0, 1]], [[0, 1, 2]] * copies
    desc = _Desc([(p, t)], [(_translate(0, 0, 0), 0)])
    o = np.array([[0.0, 2.0, -0.2]], np.float32)
    d = np.array([[0.0, -1.0, 0.0]], np.float32)
    sh = ctx.upload_scene(desc)
    sd = ctx.upload_scene(desc, device_build=True)
    assert (sd.binary_nodes, sd.blas_nodes4, sd.blas_bound, sd.levels) == \
        (sh.binary_nodes, sh.blas_nodes4, sh.blas_bound, sh.levels)
    tt, inst, prim, _ = ctx.closest_hit(o, d)
    assert inst[0] == 0 and prim[0] < copies and abs(tt[0] - 2.0) < 1e-6


def test_bad_index_rejected(ctx, cornell):
    p, t = _quad(0.0)
    t = [[0, 1, 2], [0, 2, 7]]  # vertex 7 of a 4-vertex mesh
    desc = _Desc([(p, t)], [(_translate(0, 0, 0), 0)])
    with pytest.raises(pa.PicoError) as e:
        ctx.upload_scene(desc, device_build=True)
    assert e.value.code == pa.PICO_ERR_ARG
    with pytest.raises(pa.PicoError) as e:
        ctx.closest_hit(np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32))
    assert e.value.code == pa.PICO_ERR_STATE
    ctx.upload_scene(cornell, device_build=True)  # the context still works
    ctx.closest_hit(np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32))
