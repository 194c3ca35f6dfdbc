"""A model file through the reference's import path (parse_node: every mesh
a Dielectric or a Light instance), rendered on the device (host and device
BVH builds, and with the -Skybox / -Sun options set) against the oracle on
the same description."""
import numpy as np
import pytest

import oracle
import pico_amd as pa
from test_scene_obj import room

pytestmark = pytest.mark.gpu


def _compare(ctx, scene, device_build=False):
    ctx.upload_scene(scene, device_build=device_build)
    cam = pa.Camera((0.0, 0.0, -3.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 1.0, 100.0)
    p = pa.params(48, 48, 8, 5, frame_seed=4)
    gp, gc, _, _ = ctx.render(cam, p)
    cp, cc, _, _ = oracle.CpuScene(scene).render(cam, p, rng=oracle.RNG_PCG)
    assert (gc == cc).mean() >= 0.999
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.98, (rel < 1e-3).mean()
    return gp


@pytest.mark.parametrize("device_build", [False, True])
def test_obj_room_renders_like_the_oracle(ctx, tmp_path, device_build):
    gp = _compare(ctx, pa.Scene.load_obj(room(tmp_path)), device_build)
    assert gp.mean() > 0.005  # the lamp lights the room


def test_obj_room_with_sky_and_sun_options(ctx, tmp_path):
    scene = pa.Scene.load_obj(room(tmp_path))
    rng = np.random.default_rng(5)
    scene.set_sky_image(rng.uniform(0, 2, (1, 8, 16, 4)).astype(np.float32))
    scene.set_sun((0.2, -0.4, 1.0), (3.0, 3.0, 2.5))
    _compare(ctx, scene)
