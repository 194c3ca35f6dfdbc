"""The allocation-free relMSE entry point (pico_hip_relmse_noalloc_d) against
pico_hip_relmse_d, through the C ABI: the same bits on two calls in a row,
the second on a smaller image (the static scratch keeps the first call's
parts), and both against an FP64 numpy sum."""
import numpy as np
import pytest

import pico_amd as pa

pytestmark = pytest.mark.gpu


def _inputs(rng, n, count):
    img = rng.gamma(1.5, 0.7, (n, 3)).astype(np.float32)
    ref = (img.astype(np.float64) + rng.normal(0.0, 0.1, (n, 3))) * count
    return img, ref


def test_noalloc_relmse_matches_allocating_entry_point(ctx):
    rng = np.random.default_rng(11)
    count = 64.0
    d_out = ctx.alloc(8)
    for n in (1920 * 1080, 4099):
        img, ref = _inputs(rng, n, count)
        d_img, d_ref = ctx.alloc(12 * n), ctx.alloc(24 * n)
        d_img.upload(img)
        d_ref.upload(ref)
        d_out.upload(np.array([np.nan]))
        ctx.relmse_noalloc_d(d_img, d_ref, count, n, d_out)
        ctx.sync()
        got = float(d_out.download(np.float64, 1)[0])
        want = ctx.relmse_d(d_img, d_ref, count, n)
        assert got == want, (n, got, want)
        r = ref / count
        exact = float(((img.astype(np.float64) - r) ** 2 / (r * r + 1e-2)).mean())
        assert abs(got - exact) <= 1e-12 * exact, (n, got, exact)
        d_img.free()
        d_ref.free()


def test_noalloc_relmse_edges(ctx):
    d = ctx.alloc(64)
    d.upload(np.array([5.0]))
    ctx.relmse_noalloc_d(d, d, 2.0, 0, d)
    ctx.sync()
    assert float(d.download(np.float64, 1)[0]) == 0.0
    with pytest.raises(pa.PicoError) as e:
        ctx.relmse_noalloc_d(d, d, 2.0, (1 << 28) + 1, d)
    assert e.value.code == pa.PICO_ERR_LIMIT
