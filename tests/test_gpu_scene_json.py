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


def test_json_sun_and_sky_box_render_like_the_oracle(ctx, tmp_path):
    """The Cornell box file with a GLOBALS entry: six PNG cube faces and the
    sun shining in through the open front."""
    import json

    from test_image_io import png_bytes
    path = cornell_json(tmp_path)
    doc = json.loads(path.read_text())
    rng = np.random.default_rng(11)
    names = ["px", "nx", "py", "ny", "pz", "nz"]
    for i, n in enumerate(names):
        (tmp_path / f"{n}.png").write_bytes(png_bytes(rng.integers(0, 256, (16, 16, 4)), 6, 8, seed=i))
    doc["GLOBALS"] = {"Settings": {"Skybox": [f"{n}.png" for n in names], "SunDirection": [0.2, -0.3, 1.0],
                                   "SunColour": [4.0, 4.0, 3.0]}}
    path.write_text(json.dumps(doc))
    scene = pa.Scene.load_json(path)
    assert scene.desc.use_sun and scene.desc.sky_image
    ctx.upload_scene(scene)
    cam = scene.cameras()["front"]
    p = pa.params(48, 48, 8, 5, frame_seed=3)
    gp, gc, _, _ = ctx.render(cam, p)
    cp, cc, _, _ = oracle.CpuScene(scene).render(cam, p, rng=oracle.RNG_PCG)
    assert (gc == cc).mean() >= 0.999
    rel = (np.abs(gp - cp) / (np.abs(cp) + 1e-3)).max(axis=1)
    assert (rel < 1e-3).mean() >= 0.98, (rel < 1e-3).mean()
