"""A scene loaded from the reference's JSON format, rendered on the device
(host and device BVH builds) against the oracle on the same description."""
import numpy as np
import pytest

import oracle
import pico_amd as pa
from test_scene_json import cornell_json

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("device_build", [False, True])
def test_json_cornell_renders_like_the_oracle(ctx, tmp_path, device_build):
    scene = pa.Scene.load_json(cornell_json(tmp_path))
    ctx.upload_scene(scene, device_build=device_build)
    cam = scene.cameras()["front"]
    w, h = 48, 48
    p = pa.params(w, h, 8, 5, frame_seed=3)
    gp, gc, _, _ = ctx.render(cam, p)
    cp, cc, _, _ = oracle.CpuScene(scene).render(cam, p, rng=oracle.RNG_PCG)
    assert np.array_equal(gc, cc)
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.98, (rel < 1e-3).mean()
    assert gp.mean() > 0.01  # the light reaches the camera
