"""Sky images of JSON scenes (GLOBALS/Skybox): PNG and Radiance RGBE files
decoded the way the reference's stb_image decodes them (stbi_load for cube
faces, stbi_loadf for an equirectangular image), and the sun keys.  Files are
written here from known pixels (zlib for the PNG data), so every case checks
the decoder against the pixels it encodes and, where the compiled reference
subset is present, against stb_image itself.  Host only."""
import ctypes as C
import json
import os
import struct
import zlib

import numpy as np
import pytest

import oracle
import pico_amd as pa

QUAD = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nf 1 2 3 4\n"
MAT = {"Grey": {"Type": "Constant", "Albedo": [0.5, 0.5, 0.5], "Roughness": 0.5}}
REF_SKYBOX = "/root/reference/Skybox"


# ------------------------------------------------------------- writers ----

def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa_, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa_ <= pb and pa_ <= pc else (b if pb <= pc else c)


def _row_bytes(row, depth):
    """One row of samples [pw * n] at `depth` bits, packed MSB first."""
    if depth == 16:
        return row.astype(">u2").tobytes()
    if depth == 8:
        return row.astype(np.uint8).tobytes()
    bits = ((row[:, None].astype(np.uint16) >> np.arange(depth - 1, -1, -1)) & 1).astype(np.uint8).reshape(-1)
    return np.packbits(bits).tobytes()


def _filter(rows, bpp, kinds):
    out, prev = [], bytes(len(rows[0]))
    for r, k in zip(rows, kinds):
        f = bytearray(len(r))
        for i in range(len(r)):
            a = r[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][k]
            f[i] = (r[i] - pred) & 255
        out.append(bytes([k]) + bytes(f))
        prev = r
    return out


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def png_bytes(samples, color, depth, seed=0, interlace=False, plte=None, trns=None, strategy=zlib.Z_DEFAULT_STRATEGY,
              level=6, idat_split=1):
    """A PNG of samples [h, w, n] (palette indices for colour type 3) with
    random per-row filters."""
    h, w, n = samples.shape
    rng = np.random.default_rng(seed)
    bpp = max(1, n * depth // 8)
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [_row_bytes(r.reshape(-1), depth) for r in sub]
        raw += b"".join(_filter(rows, bpp, rng.integers(0, 5, len(rows))))
    co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
    data = co.compress(raw) + co.flush()
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 1 if interlace else 0))
    out += _chunk(b"tEXt", b"Comment\x00written by the test")
    if plte is not None:
        out += _chunk(b"PLTE", np.asarray(plte, np.uint8).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    step = max(1, -(-len(data) // idat_split))
    for i in range(0, len(data), step):
        out += _chunk(b"IDAT", data[i:i + step])
    return out + _chunk(b"IEND", b"")


def rgbe_bytes(rgbe, rle, header=b"#?RADIANCE\n"):
    """A Radiance file of RGBE texels [h, w, 4]: flat, or run-length
    scanlines mixing runs and literal dumps."""
    h, w, _ = rgbe.shape
    out = header + b"# written by the test\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n" + f"-Y {h} +X {w}\n".encode()
    if not rle:
        return out + rgbe.astype(np.uint8).tobytes()
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for k in range(4):
            ch = rgbe[y, :, k].astype(np.uint8)
            i = 0
            while i < w:
                j = i
                while j < w and j - i < 127 and ch[j] == ch[i]:
                    j += 1
                if j - i >= 3:
                    out += bytes([128 + j - i, ch[i]])
                    i = j
                else:
                    j = min(w, i + 1 + int(ch[i] % 5))
                    out += bytes([j - i]) + ch[i:j].tobytes()
                    i = j
    return out


# -------------------------------------------------------------- loading ----

def _load(tmp_path, skybox, files, sun=None):
    for name, data in files.items():
        (tmp_path / name).write_bytes(data)
    (tmp_path / "quad.obj").write_text(QUAD)
    g = {"Skybox": skybox}
    if sun:
        g.update(sun)
    doc = {"GLOBALS": {"Settings": g}, "MESH": {"Quad": {"Path": "quad.obj"}}, "MATERIALS": MAT,
           "INSTANCE": {"floor": {"Asset": "Quad", "BSRDF": "Diffuse", "Material": "Grey"}}}
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(doc))
    return pa.Scene.load_json(path)


def sky_texels(scene):
    d = scene.desc
    assert d.sky_image
    k = pa.SkyImage.from_address(d.sky_image)
    n = k.width * k.height * k.faces * 4
    ct = C.c_float if k.format == pa.TEXEL_RGBA32F else C.c_uint8
    t = np.ctypeslib.as_array(C.cast(k.texels, C.POINTER(ct)), (n,)).copy()
    return t.reshape(k.faces, k.height, k.width, 4)


def decode8(tmp_path, data):
    """The library's stbi_load of one image (all six cube faces the same file)."""
    return sky_texels(_load(tmp_path, ["a.png"] * 6, {"a.png": data}))[0]


def decodef(tmp_path, data, name="a.hdr"):
    return sky_texels(_load(tmp_path, name, {name: data}))[0]


def _check_ref(tmp_path, name, got, want_float):
    if not oracle.ref_available():
        return
    ref = oracle.ref_image(tmp_path / name, want_float)
    assert ref is not None
    if want_float:
        np.testing.assert_array_equal(got, ref)
    else:
        assert np.array_equal(got, ref)


# ---------------------------------------------------------------- PNG ----

SCALE = {1: 0xFF, 2: 0x55, 4: 0x11, 8: 1}


def _expect(samples, color, depth):
    """stbi_load's RGBA8 of the encoded samples."""
    s = samples.astype(np.int64)
    s = s >> 8 if depth == 16 else s * SCALE[depth]
    h, w, n = s.shape
    out = np.empty((h, w, 4), np.int64)
    if color in (0, 4):
        out[..., :3] = s[..., :1]
    else:
        out[..., :3] = s[..., :3]
    out[..., 3] = s[..., n - 1] if color in (4, 6) else 255
    return out.astype(np.uint8)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("color,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (4, 8), (4, 16),
                                         (6, 8), (6, 16)])
def test_png_colour_types_and_depths(tmp_path, color, depth, interlace):
    n = {0: 1, 2: 3, 4: 2, 6: 4}[color]
    rng = np.random.default_rng(color * 100 + depth)
    h, w = 13, 21
    samples = rng.integers(0, 1 << depth, (h, w, n))
    data = png_bytes(samples, color, depth, seed=depth, interlace=interlace)
    got = decode8(tmp_path, data)
    assert np.array_equal(got, _expect(samples, color, depth))
    _check_ref(tmp_path, "a.png", got, False)


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_png_palette_and_alpha(tmp_path, depth):
    rng = np.random.default_rng(depth)
    n_pal = 1 << depth
    pal = rng.integers(0, 256, (n_pal, 3))
    alpha = rng.integers(0, 256, max(1, n_pal // 2))  # a tRNS shorter than the palette
    idx = rng.integers(0, n_pal, (9, 17, 1))
    data = png_bytes(idx, 3, depth, plte=pal, trns=alpha.astype(np.uint8).tobytes())
    got = decode8(tmp_path, data)
    a = np.full(n_pal, 255)
    a[:len(alpha)] = alpha
    want = np.concatenate([pal[idx[..., 0]], a[idx[..., 0]][..., None]], axis=2).astype(np.uint8)
    assert np.array_equal(got, want)
    _check_ref(tmp_path, "a.png", got, False)


@pytest.mark.parametrize("color,depth", [(0, 8), (0, 2), (2, 8), (2, 16), (0, 16)])
def test_png_colour_key(tmp_path, color, depth):
    n = 3 if color == 2 else 1
    rng = np.random.default_rng(depth + color)
    samples = rng.integers(0, 3, (11, 10, n)) * ((1 << depth) - 1) // 2
    key = samples[0, 0]
    trns = b"".join(struct.pack(">H", int(k)) for k in key)
    got = decode8(tmp_path, png_bytes(samples, color, depth, trns=trns))
    want = _expect(samples, color, depth)
    want[..., 3] = np.where((samples == key).all(axis=2), 0, 255)
    assert np.array_equal(got, want)
    _check_ref(tmp_path, "a.png", got, False)


@pytest.mark.parametrize("strategy,level,split", [(zlib.Z_DEFAULT_STRATEGY, 0, 1), (zlib.Z_FIXED, 6, 1),
                                                  (zlib.Z_HUFFMAN_ONLY, 9, 3), (zlib.Z_RLE, 6, 7),
                                                  (zlib.Z_DEFAULT_STRATEGY, 9, 64)])
def test_png_deflate_block_kinds(tmp_path, strategy, level, split):
    """Stored, fixed-code and dynamic-code blocks, long matches and IDAT data
    split over many chunks."""
    rng = np.random.default_rng(level + split)
    h, w = 70, 90
    base = rng.integers(0, 256, (h, w, 4))
    base[10:40] = base[10]          # long repeats (matches up to 258 bytes)
    base[:, 50:] = base[:, 50:51]
    data = png_bytes(base, 6, 8, seed=level, strategy=strategy, level=level, idat_split=split)
    got = decode8(tmp_path, data)
    assert np.array_equal(got, base.astype(np.uint8))
    _check_ref(tmp_path, "a.png", got, False)


def test_png_as_float_is_gamma_linearised(tmp_path):
    """stbi_loadf of an 8-bit image: powf(v / 255, 2.2) for colour, v / 255
    for alpha."""
    rng = np.random.default_rng(5)
    px = rng.integers(0, 256, (6, 256 // 6 + 1, 4))
    px.reshape(-1, 4)[:256, 0] = np.arange(256)  # every value
    got = decodef(tmp_path, png_bytes(px, 6, 8), "a.png")
    want = (px / np.float32(255)).astype(np.float32)
    want[..., :3] = np.power(want[..., :3], np.float32(2.2))
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=0)
    _check_ref(tmp_path, "a.png", got, True)


# --------------------------------------------------------------- RGBE ----

@pytest.mark.parametrize("w,rle", [(5, False), (40, False), (40, True), (300, True)])
def test_rgbe(tmp_path, w, rle):
    rng = np.random.default_rng(w)
    h = 7
    rgbe = rng.integers(0, 256, (h, w, 4))
    rgbe[..., 3] = rng.integers(100, 160, (h, w))
    rgbe[2, :, 3] = 0                    # exponent 0: black
    rgbe[3, 2:30] = rgbe[3, 2]           # runs
    rgbe[0, 0, 0] = 7                    # a flat file must not start like a run-length scanline (2, 2)
    data = rgbe_bytes(rgbe, rle and w >= 8, header=b"#?RGBE\n" if w == 40 else b"#?RADIANCE\n")
    got = decodef(tmp_path, data)
    m = rgbe[..., :3].astype(np.float64) * np.ldexp(1.0, rgbe[..., 3:4] - 136)
    m[rgbe[..., 3] == 0] = 0.0
    want = np.concatenate([m, np.ones((h, w, 1))], axis=2).astype(np.float32)
    assert np.array_equal(got, want)
    _check_ref(tmp_path, "a.hdr", got, True)


# -------------------------------------------------------- scene level ----

def test_cube_faces_and_sun(tmp_path):
    rng = np.random.default_rng(9)
    faces = rng.integers(0, 256, (6, 8, 12, 4))
    files = {f"f{i}.png": png_bytes(faces[i], 6, 8, seed=i) for i in range(6)}
    sun = {"SunDirection": [1.0, -2.0, 0.5], "SunColour": [3.0, 2.0, 1.0]}
    s = _load(tmp_path, [f"f{i}.png" for i in range(6)], files, sun)
    assert np.array_equal(sky_texels(s), faces.astype(np.uint8))
    k = pa.SkyImage.from_address(s.desc.sky_image)
    assert (k.faces, k.format, k.width, k.height) == (6, pa.TEXEL_RGBA8, 12, 8)
    d = np.array([1.0, -2.0, 0.5], np.float32)
    want = d * (np.float32(1) / np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
    assert s.desc.use_sun == 1 and list(s.desc.sun_colour) == [3.0, 2.0, 1.0]
    assert np.array_equal(np.array(s.desc.sun_direction, np.float32), want.astype(np.float32))


def test_sun_needs_both_keys_and_last_globals_entry_wins(tmp_path):
    (tmp_path / "quad.obj").write_text(QUAD)
    (tmp_path / "sky.hdr").write_bytes(rgbe_bytes(np.full((2, 3, 4), 130), False))
    doc = {"GLOBALS": {"A": {"Skybox": "sky.hdr", "SunDirection": [0, -1, 0], "SunColour": [1, 1, 1]},
                       "B": {"SunDirection": [0, -1, 1]}},
           "MESH": {"Quad": {"Path": "quad.obj"}}, "MATERIALS": MAT,
           "INSTANCE": {"floor": {"Asset": "Quad", "BSRDF": "Diffuse", "Material": "Grey"}}}
    (tmp_path / "scene.json").write_text(json.dumps(doc))
    s = pa.Scene.load_json(tmp_path / "scene.json")
    # "B" comes last: no sky box (black) and a direction without a colour (no sun)
    assert not s.desc.sky_image and s.desc.use_sun == 0 and list(s.desc.sky_rgb) == [0, 0, 0]


@pytest.mark.parametrize("case,needle", [
    ("notpng", "not a PNG"), ("truncated", "ends"), ("badzlib", "zlib"), ("sizes", "differ in size"),
    ("five", "six faces"), ("missing", "cannot read"), ("hdrface", "Radiance"), ("badrgbe", "FORMAT")])
def test_sky_errors(tmp_path, case, needle):
    good = png_bytes(np.zeros((4, 4, 4), int), 6, 8)
    files = {"a.png": good, "b.png": png_bytes(np.zeros((5, 4, 4), int), 6, 8)}
    sky = ["a.png"] * 6
    if case == "notpng":
        files["a.png"] = b"\xff\xd8\xff\xe0 a JPEG"
    elif case == "truncated":
        files["a.png"] = good[:40]
    elif case == "badzlib":
        i = good.index(b"IDAT") + 4
        files["a.png"] = good[:i] + b"\x00\x00" + good[i + 2:]
    elif case == "sizes":
        sky = ["a.png"] * 5 + ["b.png"]
    elif case == "five":
        sky = ["a.png"] * 5
    elif case == "missing":
        sky = ["a.png"] * 5 + ["nope.png"]
    elif case == "hdrface":
        files["a.png"] = rgbe_bytes(np.full((2, 3, 4), 130), False)
    elif case == "badrgbe":
        files["a.hdr"] = b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x00\x00\x00\x00"
        sky = "a.hdr"
    with pytest.raises(pa.PicoError) as e:
        _load(tmp_path, sky, files)
    assert needle in str(e.value), str(e.value)


@pytest.mark.skipif(not os.path.isdir(REF_SKYBOX) or not oracle.ref_available(),
                    reason="the reference's sky box and compiled subset are not here")
def test_reference_skybox_faces(tmp_path):
    """The reference's own sky box (six 1024^2 RGBA PNG faces, in the order
    px, nx, py, ny, pz, nz) decodes to stb_image's texels."""
    names = ["px", "nx", "py", "ny", "pz", "nz"]
    doc = {"GLOBALS": {"Settings": {"Skybox": [os.path.join(REF_SKYBOX, f"{n}.png") for n in names]}},
           "MESH": {"Quad": {"Path": "quad.obj"}}, "MATERIALS": MAT,
           "INSTANCE": {"floor": {"Asset": "Quad", "BSRDF": "Diffuse", "Material": "Grey"}}}
    (tmp_path / "quad.obj").write_text(QUAD)
    (tmp_path / "scene.json").write_text(json.dumps(doc))
    t = sky_texels(pa.Scene.load_json(tmp_path / "scene.json"))
    assert t.shape == (6, 1024, 1024, 4)
    for i, n in enumerate(names):
        assert np.array_equal(t[i], oracle.ref_image(os.path.join(REF_SKYBOX, f"{n}.png"))), n
