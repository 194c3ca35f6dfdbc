"""Test scenes with the sun light and a sky image (shared by the oracle and
device tests)."""
import numpy as np

import pico_amd as pa


def sun_cube_sky(microfacet=pa.MICROFACET_BECKMANN_ALL):
    """The small instanced field lit by the sun from above, under a cube-map
    sky of random RGBA8 faces."""
    s = pa.Scene.instanced(7, 4, 48, 200, 2000, 12.0, microfacet)
    faces = np.random.default_rng(3).integers(0, 256, (6, 16, 12, 4), dtype=np.uint8)
    return s.set_sun((0.3, -0.9, 0.2), (3.0, 2.5, 2.0)).set_sky_image(faces)


def sun_hdri_sky(microfacet=pa.MICROFACET_BECKMANN_ALL):
    """The Cornell box with sunlight entering through its open front, under
    an equirectangular float sky."""
    s = pa.Scene.cornell(1, microfacet)
    d = np.array([0.2, -0.3, 1.0])
    sky = (np.random.default_rng(4).random((1, 24, 48, 4)) * 2).astype(np.float32)
    return s.set_sun(tuple(d / np.linalg.norm(d)), (4.0, 4.0, 3.0)).set_sky_image(sky)


def sky_dirs(n, seed):
    """Random unit directions plus the axes and diagonals (face ties)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [1, 0, 1],
                     [0, 1, 1], [1, 1, 1], [-1, -1, -1], [1, -1, 0], [-1, 0, 1], [0, -1, -1]], float)
    d = np.concatenate([d, axes])
    return np.ascontiguousarray((d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))


def shapes_sun_sky(microfacet=pa.MICROFACET_BECKMANN_ALL):
    """Every feature at once: the built-in sphere and cube (also as lights),
    Fresnel and Transparent glass with the index-of-refraction stack, the sun
    shining in through the open front and a cube-map sky."""
    s = pa.Scene.shapes(1, microfacet)
    d = np.array([-0.3, -0.4, 1.0])
    faces = np.random.default_rng(5).integers(0, 256, (6, 8, 8, 4), dtype=np.uint8)
    return s.set_sun(tuple(d / np.linalg.norm(d)), (2.0, 2.0, 1.5)).set_sky_image(faces)
