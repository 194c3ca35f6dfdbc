"""The .hdr writer (the file half of Scene::render_scene_to_file,
Source/Core/Scene.cpp:236-246) against the RGBE definition and against the
reference's own stb writer and reader (compiled reference subset)."""
import numpy as np
import pytest

import oracle
import pico_amd as pa


def _image(w, h, seed=0):
    rng = np.random.default_rng(seed)
    img = rng.gamma(1.0, 2.0, (w * h, 3)).astype(np.float32)
    img[0] = 0.0                      # black pixel -> (0, 0, 0, 0)
    img[1] = (1e-40, 0.0, 0.0)        # below the RGBE floor
    img[2] = (1000.0, 0.5, 1e-3)      # wide range in one pixel
    return img


def _read_flat_rgbe(path, w, h):
    data = open(path, "rb").read()
    head, _, body = data.partition(b"\n\n")
    assert head.startswith(b"#?RADIANCE") and b"FORMAT=32-bit_rle_rgbe" in head
    res, _, px = body.partition(b"\n")
    assert res == f"-Y {h} +X {w}".encode()
    q = np.frombuffer(px, np.uint8).reshape(-1, 4).astype(np.float64)
    scale = np.where(q[:, 3] > 0, np.ldexp(1.0, (q[:, 3] - 136).astype(int)), 0.0)
    return q[:, :3] * scale[:, None]


def test_rgbe_definition_and_reversal(tmp_path):
    w, h = 13, 7
    img = _image(w, h)
    path = str(tmp_path / "a.hdr")
    pa.write_hdr(path, img, w, h, reverse=True)
    dec = _read_flat_rgbe(path, w, h)
    exp = img[::-1].astype(np.float64)
    # truncation to 8 bits of mantissa under the brightest channel's exponent
    # e (v = m 2^e, m in [0.5, 1)): error below 2^(e - 8) <= 2^-7 v
    m = exp.max(axis=1, keepdims=True)
    assert np.all(np.abs(dec - exp) <= m * 2.0 ** -7 + 1e-30)
    assert np.all(dec[-1] == 0) and np.all(dec[-2] == 0)


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference subset not built")
def test_same_pixels_as_the_reference_writer(tmp_path):
    w, h = 40, 23
    img = _image(w, h, 1)
    ours, theirs = str(tmp_path / "ours.hdr"), str(tmp_path / "ref.hdr")
    pa.write_hdr(ours, img, w, h, reverse=True)
    oracle.ref_write_hdr(theirs, np.ascontiguousarray(img[::-1]), w, h)
    a, b = oracle.ref_load_hdr(ours, w, h), oracle.ref_load_hdr(theirs, w, h)
    np.testing.assert_array_equal(a, b)
