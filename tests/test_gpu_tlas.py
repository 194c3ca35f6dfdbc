"""Device TLAS rebuild (new instance transforms, linear BVH built on the GPU):
closest hits and rendered images against the host-built TLAS of the same
transforms and against the oracle, and the rejection of bad transforms with
the scene left as it was."""
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


def _transforms(scene):
    d = scene.desc
    return np.ctypeslib.as_array(C.cast(d.transforms, C.POINTER(C.c_float)), (d.n_instances * 16,)).copy()


def _perturbed(scene, seed, angle=0.3, shift=0.5):
    """Every instance turned about its world-space origin (a random axis,
    up to `angle` radians) and moved by up to `shift` per axis."""
    xf = _transforms(scene).astype(np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)  # row-major matrices
    rng = np.random.default_rng(seed)
    out = np.empty_like(xf)
    for i, m in enumerate(xf):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        a = rng.uniform(-angle, angle)
        k = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        r = np.eye(3) + np.sin(a) * k + (1 - np.cos(a)) * (k @ k)
        c = m[:3, 3]
        t = np.eye(4)
        t[:3, :3] = r
        t[:3, 3] = c - r @ c + rng.uniform(-shift, shift, 3)
        out[i] = t @ m
        out[i, 3] = (0.0, 0.0, 0.0, 1.0)
    return np.ascontiguousarray(out.transpose(0, 2, 1).reshape(-1, 16), dtype=np.float32)


class _Moved:
    """A scene description with other instance transforms (and optionally
    fewer instances and no lights)."""

    def __init__(self, scene, transforms, n_instances=None):
        self.scene = scene  # keeps the arrays alive
        self.xf = np.ascontiguousarray(transforms, np.float32)
        self.desc = pa.SceneDesc.from_buffer_copy(scene.desc)
        self.desc.transforms = self.xf.ctypes.data
        if n_instances is not None:
            self.desc.n_instances = n_instances
            self.desc.n_lights = 0

    def camera(self, aspect):
        return self.scene.camera(aspect)


def _random_rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def _same_hits(a, b, min_frac=0.9999):
    """Closest hits of two traversals: the same instance and primitive for all
    but tie-breaking rays, and there the same distance and barycentrics, bit
    for bit (the leaf arithmetic does not depend on the tree)."""
    same = (a[1] == b[1]) & (a[2] == b[2])
    assert same.mean() >= min_frac, same.mean()
    assert np.array_equal(a[0][same], b[0][same])
    assert np.array_equal(a[3][same], b[3][same])
    return same


def test_same_transforms_give_same_hits_and_image(ctx, instanced):
    ctx.upload_scene(instanced)
    o, d = _random_rays(50000, 5, 0.0, 100.0)
    before = ctx.closest_hit(o, d)
    cam = instanced.camera(16 / 9)
    p = pa.params(192, 108, 2, 6, frame_seed=4)
    img0, cnt0, _, _ = ctx.render(cam, p)
    st = ctx.set_instance_transforms(_transforms(instanced))
    n = instanced.desc.n_instances
    print(f"[tlas] rebuild of {n} instances: {st.device_ms:.3f} ms, depth {st.binary_depth}, "
          f"{st.wide_nodes} wide nodes, stack bound {st.stack_bound}", flush=True)
    assert 0 < st.wide_nodes <= n - 1 and st.stack_bound <= 64
    after = ctx.closest_hit(o, d)
    _same_hits(before, after)
    img1, cnt1, _, _ = ctx.render(cam, p)
    assert np.array_equal(cnt0, cnt1)
    eq = np.all(img0 == img1, axis=1)
    assert eq.mean() >= 0.999, eq.mean()


def test_rebuild_matches_host_build_and_oracle(ctx, instanced):
    xf = _perturbed(instanced, 7)
    moved = _Moved(instanced, xf)
    o, d = _random_rays(50000, 6, 0.0, 100.0)
    ctx.upload_scene(moved)  # host-built TLAS of the new transforms
    host = ctx.closest_hit(o, d)
    ctx.upload_scene(instanced)
    ctx.set_instance_transforms(xf)
    dev = ctx.closest_hit(o, d)
    _same_hits(host, dev)
    ct, ci, cp, *_ = oracle.CpuScene(moved).closest_hit(o, d)
    assert (ci != 0xffffffff).mean() > 0.3
    agree = (dev[1] == ci) & (dev[2] == cp)
    assert agree.mean() >= 0.9999, agree.mean()
    m = agree & np.isfinite(ct)
    assert np.isclose(dev[0][m], ct[m], rtol=1e-4, atol=1e-4).mean() >= 0.9999


def test_repeated_rebuilds_from_device_memory(ctx, instanced):
    """Several frames of animation through the `_d` entry point, each checked
    against the host build of the same transforms."""
    ctx.upload_scene(instanced)
    o, d = _random_rays(20000, 8, 0.0, 100.0)
    n = instanced.desc.n_instances
    buf = ctx.alloc(64 * n)
    for frame in range(3):
        xf = _perturbed(instanced, 100 + frame, angle=0.1 * (frame + 1), shift=2.0)
        buf.upload(xf)
        ctx.set_instance_transforms_d(buf)
        dev = ctx.closest_hit(o, d)
        moved = _Moved(instanced, xf)
        ctx.upload_scene(moved)
        host = ctx.closest_hit(o, d)
        _same_hits(host, dev)
        ctx.upload_scene(instanced)  # next frame starts from the original upload
    buf.free()


def test_render_after_rebuild_matches_oracle(ctx, cornell):
    xf = _perturbed(cornell, 9, angle=0.05, shift=0.02)
    moved = _Moved(cornell, xf)
    ctx.upload_scene(cornell)
    ctx.set_instance_transforms(xf)
    w, h = 48, 48
    cam = cornell.camera(1.0)
    p = pa.params(w, h, 8, 5, frame_seed=2)
    gp, gc, _, _ = ctx.render(cam, p)
    cp, cc, _, _ = oracle.CpuScene(moved).render(cam, p, rng=oracle.RNG_PCG)
    assert np.array_equal(gc, cc)
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.98, (rel < 1e-3).mean()


def test_single_instance(ctx, cornell):
    xf = _transforms(cornell)[:16]
    one = _Moved(cornell, xf, n_instances=1)
    ctx.upload_scene(one)
    new = _perturbed(one, 10, angle=0.4, shift=0.1)
    st = ctx.set_instance_transforms(new)
    assert st.wide_nodes == 1
    o, d = _random_rays(20000, 11, -0.9, 0.9)
    dev = ctx.closest_hit(o, d)
    ct, ci, cp, *_ = oracle.CpuScene(_Moved(cornell, new, n_instances=1)).closest_hit(o, d)
    assert (ci != 0xffffffff).mean() > 0.05
    agree = (dev[1] == ci) & (dev[2] == cp)
    assert agree.mean() >= 0.9999, agree.mean()


@pytest.mark.parametrize("kind", ["singular", "projective", "nan"])
def test_bad_transform_rejected_scene_unchanged(ctx, cornell, kind):
    ctx.upload_scene(cornell)
    o, d = _random_rays(5000, 12, -0.9, 0.9)
    before = ctx.closest_hit(o, d)
    xf = _perturbed(cornell, 13).reshape(-1, 16)
    if kind == "singular":
        xf[3, 0:3] = 0.0  # first column of instance 3 zero
    elif kind == "projective":
        xf[2, 3] = 0.5    # bottom row of instance 2 not (0, 0, 0, 1)
    else:
        xf[1, 5] = np.nan
    with pytest.raises(pa.PicoError) as e:
        ctx.set_instance_transforms(xf)
    assert e.value.code == pa.PICO_ERR_ARG
    after = ctx.closest_hit(o, d)
    for a, b in zip(before, after):
        assert np.array_equal(a, b, equal_nan=True)


def test_many_instances(ctx):
    """65536 small instances: the rebuilt TLAS against the host build."""
    scene = pa.Scene.instanced(3, 16, 65536, 20, 200, 400.0, pa.MICROFACET_GGX)
    xf = _perturbed(scene, 14, angle=1.0, shift=3.0)
    moved = _Moved(scene, xf)
    o, d = _random_rays(50000, 15, 0.0, 400.0)
    ctx.upload_scene(moved)
    host = ctx.closest_hit(o, d)
    ctx.upload_scene(scene)
    st = ctx.set_instance_transforms(xf)
    print(f"[tlas] rebuild of 65536 instances: {st.device_ms:.3f} ms, depth {st.binary_depth}, "
          f"{st.wide_nodes} wide nodes, stack bound {st.stack_bound}", flush=True)
    dev = ctx.closest_hit(o, d)
    assert (host[1] != 0xffffffff).mean() > 0.3
    _same_hits(host, dev)
