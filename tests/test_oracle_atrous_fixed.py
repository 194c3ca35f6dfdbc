"""The fixed-width a-trous oracle (labelled mode, no reference counterpart)
pinned by closed forms: a constant image is a fixed point; with every
edge-stopping term off it is the clamped, dilated B3-spline blur computed
here by brute force; a single bright pixel under the colour term alone
keeps the weight the definition gives it; edges in normals stop the blur;
the colour width halves from one level to the next."""
import numpy as np

import oracle


def _blur_b3(img, w, h, levels):
    """Brute-force separable-in-definition B3 blur with clamped taps."""
    hb = np.array([1, 4, 6, 4, 1], np.float64) / 16
    a = img.reshape(h, w, 3).astype(np.float64)
    for l in range(levels):
        s = 1 << l
        out = np.zeros_like(a)
        for y in range(h):
            for x in range(w):
                acc = np.zeros(3)
                for v in range(5):
                    for u in range(5):
                        yy = min(max(y + (v - 2) * s, 0), h - 1)
                        xx = min(max(x + (u - 2) * s, 0), w - 1)
                        acc += hb[u] * hb[v] * a[yy, xx]
                out[y, x] = acc  # the weights sum to 1
        a = out
    return a.reshape(-1, 3)


def _gbuf(rng, n):
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pos = (rng.normal(size=(n, 3)) * 5).astype(np.float32)
    alb = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    return nrm, pos, alb


def test_constant_image_is_a_fixed_point():
    rng = np.random.default_rng(1)
    w, h = 23, 17
    nrm, pos, alb = _gbuf(rng, w * h)
    rgb = np.tile(np.array([[0.3, 1.5, 7.0]], np.float32), (w * h, 1))
    out = oracle.atrous_fixed(rgb, nrm, pos, alb, w, h, 5, (0.3, 0.05, 0.1, 0.5))
    np.testing.assert_allclose(out, rgb, rtol=1e-6)


def test_all_terms_off_is_the_dilated_b3_blur():
    rng = np.random.default_rng(2)
    w, h = 19, 13
    nrm, pos, alb = _gbuf(rng, w * h)
    rgb = rng.gamma(1.0, 1.0, (w * h, 3)).astype(np.float32)
    out = oracle.atrous_fixed(rgb, nrm, pos, alb, w, h, 3, (0.0, 0.0, 0.0, 0.0))
    np.testing.assert_allclose(out, _blur_b3(rgb, w, h, 3), rtol=2e-6, atol=1e-7)


def test_single_bright_pixel_under_the_colour_term():
    # one row: every vertical tap clamps onto the row, so the kernel is the
    # 1-D B3 spline; the centre keeps h(0) against (1 - h(0)) exp(-3 / sc^2)
    w, h, sc = 5, 1, 1.3
    rgb = np.zeros((w, 3), np.float32)
    rgb[2] = 1.0
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (w, 1))
    pos = np.tile(np.array([[0, 0, 4]], np.float32), (w, 1))
    alb = np.full((w, 3), 0.5, np.float32)
    out = oracle.atrous_fixed(rgb, nrm, pos, alb, w, h, 1, (0.0, 0.0, 0.0, sc))
    h0 = 6 / 16
    expect = h0 / (h0 + (1 - h0) * np.exp(-3 / sc ** 2))
    np.testing.assert_allclose(out[2], expect, rtol=1e-6)


def test_normal_edge_stops_the_blur():
    w, h = 16, 8
    rgb = np.zeros((h, w, 3), np.float32)
    rgb[:, w // 2:] = 2.0
    nrm = np.zeros((h, w, 3), np.float32)
    nrm[:, : w // 2, 2] = 1.0
    nrm[:, w // 2:, 0] = 1.0
    pos = np.tile(np.array([0, 0, 4], np.float32), (h, w, 1))
    alb = np.full((h, w, 3), 0.5, np.float32)
    out = oracle.atrous_fixed(rgb.reshape(-1, 3), nrm.reshape(-1, 3), pos.reshape(-1, 3), alb.reshape(-1, 3), w, h, 5,
                              (0.1, 0.0, 0.0, 0.0)).reshape(h, w, 3)
    # |dn|^2 = 2 across the edge: weight exp(-200), nothing crosses
    np.testing.assert_allclose(out, rgb, atol=1e-30)


def _colour_pass(img, w, h, step, sigma):
    """One pass of the colour-only edge-stopping B3 filter (Dammertz et al.
    2010): weight h(u) h(v) exp(-|c_q - c_p|^2 / sigma^2), clamped taps at
    offsets step * (-2..2), normalised by the weight sum."""
    hb = np.array([1, 4, 6, 4, 1], np.float64) / 16
    a = img.reshape(h, w, 3).astype(np.float64)
    out = np.zeros_like(a)
    for y in range(h):
        for x in range(w):
            acc, ws = np.zeros(3), 0.0
            for v in range(5):
                for u in range(5):
                    yy = min(max(y + (v - 2) * step, 0), h - 1)
                    xx = min(max(x + (u - 2) * step, 0), w - 1)
                    wt = hb[u] * hb[v] * np.exp(-np.sum((a[yy, xx] - a[y, x]) ** 2) / sigma ** 2)
                    acc += wt * a[yy, xx]
                    ws += wt
            out[y, x] = acc / ws
    return out.reshape(-1, 3)


def test_colour_width_halves_per_level():
    """Level l uses the colour width sigma / 2^l: with the other terms off,
    two levels equal one level followed by a step-2 pass at half the width
    (and clearly not a step-2 pass at the full width)."""
    rng = np.random.default_rng(11)
    w, h = 13, 9
    rgb = rng.uniform(0, 1, (w * h, 3)).astype(np.float32)
    nrm, pos, alb = _gbuf(rng, w * h)
    sigma = 0.6
    one = oracle.atrous_fixed(rgb, nrm, pos, alb, w, h, 1, (0.0, 0.0, 0.0, sigma))
    two = oracle.atrous_fixed(rgb, nrm, pos, alb, w, h, 2, (0.0, 0.0, 0.0, sigma))
    np.testing.assert_allclose(one, _colour_pass(rgb, w, h, 1, sigma), rtol=1e-5, atol=1e-6)
    halved = _colour_pass(one, w, h, 2, sigma / 2)
    np.testing.assert_allclose(two, halved, rtol=1e-5, atol=1e-6)
    assert np.abs(_colour_pass(one, w, h, 2, sigma) - halved).max() > 1e-2
