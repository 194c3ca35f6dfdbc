"""Driver of the sanitizer tests: runs the host parsers and scene code
(pico_amd/csrc/_san/libpico_host_san.so) and the oracle
(oracle/_san/libcpu_ref_san.so), both built with AddressSanitizer and
UndefinedBehaviorSanitizer, over valid and damaged inputs.  Run as a child
process with the sanitizer runtimes preloaded (tests/test_host_sanitizers.py);
a sanitizer report aborts it with a non-zero status.  Prints one line per
group of cases."""
import ctypes as C
import json
import os
import sys
import tempfile
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import pico_amd as pa  # noqa: E402  (structures only; the GPU library is not loaded)
from test_image_io import png_bytes, rgbe_bytes  # noqa: E402

HOST = C.CDLL(os.path.join(ROOT, "pico_amd", "csrc", "_san", "libpico_host_san.so"))
ORC = C.CDLL(os.path.join(ROOT, "oracle", "_san", "libcpu_ref_san.so"))
vp, u32, f32, i32, sz = C.c_void_p, C.c_uint32, C.c_float, C.c_int32, C.c_size_t
HOST.pico_scene_load_json.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, sz]
HOST.pico_scene_desc_get.argtypes = [vp, C.POINTER(pa.SceneDesc)]
HOST.pico_scene_destroy.argtypes = [vp]
HOST.pico_scene_destroy.restype = None
HOST.pico_scene_camera_count.argtypes = [vp, C.POINTER(u32)]
HOST.pico_scene_camera_get.argtypes = [vp, u32, C.c_char_p, sz, C.POINTER(pa.Camera)]
HOST.pico_scene_default_camera.argtypes = [vp, f32, C.POINTER(pa.Camera)]
HOST.pico_scene_cornell.argtypes = [u32, u32, C.POINTER(vp)]
HOST.pico_scene_instanced.argtypes = [u32, u32, u32, u32, u32, f32, u32, C.POINTER(vp)]
HOST.pico_scene_training.argtypes = [u32, u32, C.POINTER(vp)]
HOST.pico_scene_shapes.argtypes = [u32, u32, C.POINTER(vp)]
HOST.pico_scene_set_sky_image.argtypes = [vp, C.POINTER(pa.SkyImage)]
HOST.pico_write_hdr.argtypes = [C.c_char_p, u32, u32, vp, i32]

QUAD = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nf 1 2 3 4\n"
DOC = {"GLOBALS": {"S": {"SunDirection": [0.3, -1, 0.2], "SunColour": [2, 2, 2]}},
       "MESH": {"Quad": {"Path": "quad.obj"}, "Box": {"Path": "box.obj"}},
       "MATERIALS": {"Grey": {"Type": "Constant", "Albedo": [0.5, 0.5, 0.5], "Roughness": 0.5},
                     "Lamp": {"Type": "Constant", "Albedo": [1, 1, 1], "Emissive": [4, 4, 4]},
                     "Glass": {"Type": "Constant", "Albedo": [1, 1, 1], "Roughness": 0.1,
                               "IndexOfRefraction": 1.5}},
       "INSTANCE": {"floor": {"Asset": "Quad", "BSRDF": "Diffuse", "Material": "Grey"},
                    "box": {"Asset": "Box", "BSRDF": "Fresnel", "Material": "Glass", "Position": [0, 1, 0],
                            "Rotation": [0, 0.38, 0, 0.92], "Scale": [0.5, 0.5, 0.5]},
                    "ball": {"Asset": "Sphere", "BSRDF": "Light", "Material": "Lamp", "Position": [1, 2, 0]}},
       "CAMERA": {"main": {"Position": [0, 1, -4], "Direction": [0, -0.2, 1], "FarPlane": 100}}}
BOX = ("v -1 -1 -1\nv 1 -1 -1\nv 1 1 -1\nv -1 1 -1\nv -1 -1 1\nv 1 -1 1\nv 1 1 1\nv -1 1 1\n"
       "vn 0 0 -1\nvt 0 0\n"
       "f 1/1/1 2/1/1 3/1/1 4/1/1\nf 5 8 7 6\nf 1 5 6 2\nf 4 3 7 8\nf -8 -4 -1 -5\nf 2 6 7 3\n")


def load(path):
    h = vp()
    err = C.create_string_buffer(512)
    rc = HOST.pico_scene_load_json(os.fsencode(path), C.byref(h), err, 512)
    if rc == 0:
        d = pa.SceneDesc()
        assert HOST.pico_scene_desc_get(h, C.byref(d)) == 0
        n = u32()
        HOST.pico_scene_camera_count(h, C.byref(n))
        cam = pa.Camera()
        name = C.create_string_buffer(64)
        for i in range(n.value):
            HOST.pico_scene_camera_get(h, i, name, 64, C.byref(cam))
        HOST.pico_scene_default_camera(h, 1.0, C.byref(cam))
        HOST.pico_scene_destroy(h)
    else:
        assert rc == pa.PICO_ERR_ARG, rc
    return rc


def mutate(rng, data, n):
    b = bytearray(data)
    for _ in range(n):
        k = int(rng.integers(0, 5))
        i = int(rng.integers(0, max(1, len(b))))
        if k == 0 and b:
            b[i % len(b)] = int(rng.integers(0, 256))
        elif k == 1 and b:
            del b[i % len(b):i % len(b) + int(rng.integers(1, 9))]
        elif k == 2:
            b[i:i] = bytes(rng.integers(0, 256, int(rng.integers(1, 9))).astype(np.uint8))
        elif k == 3:
            b[i:i] = rng.choice([b"{", b"}", b"[", b"]", b",", b":", b'"', b"-", b"1e999", b"//", b"/*", b"\\u"])
        else:
            del b[i:]
    return bytes(b)


def scene_files(tmp):
    with open(os.path.join(tmp, "quad.obj"), "w") as f:
        f.write(QUAD)
    with open(os.path.join(tmp, "box.obj"), "w") as f:
        f.write(BOX)


def main():
    rng = np.random.default_rng(int(os.environ.get("PICO_SAN_SEED", "1")))
    n_mut = int(os.environ.get("PICO_SAN_MUTATIONS", "600"))
    with tempfile.TemporaryDirectory() as tmp:
        scene_files(tmp)
        path = os.path.join(tmp, "scene.json")
        text = json.dumps(DOC, indent=1).encode()
        with open(path, "wb") as f:
            f.write(text)
        assert load(path) == 0
        # damaged scene files
        ok = 0
        for _ in range(n_mut):
            with open(path, "wb") as f:
                f.write(mutate(rng, text, int(rng.integers(1, 6))))
            ok += load(path) == 0
        print(f"json: {n_mut} mutated scene files, {ok} accepted", flush=True)
        # damaged OBJ meshes under the valid scene file
        with open(path, "wb") as f:
            f.write(text)
        ok = 0
        for case in ["", "f 1 2 3\n", "v 1 2\nf 1 1 1\n", "v 0 0 0\nf 0 1 2\n", "v 0 0 0\nf -5 1 1\n",
                     "v 1e999 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\n", "v " + "9" * 5000 + " 0 0\n",
                     "vn 0 0 1\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//7 2//1 3//1\n", "v 0 0 0\nf 1/2/3/4 1 1 1 1\n",
                     "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3", "v nan 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"]:
            with open(os.path.join(tmp, "box.obj"), "w") as f:
                f.write(case)
            ok += load(path) == 0
        for _ in range(n_mut // 2):
            with open(os.path.join(tmp, "box.obj"), "wb") as f:
                f.write(mutate(rng, BOX.encode(), int(rng.integers(1, 6))))
            ok += load(path) == 0
        print(f"obj: {11 + n_mut // 2} damaged meshes, {ok} accepted", flush=True)
        scene_files(tmp)
        # sky images: valid PNG / RGBE files of every kind, then damaged ones
        doc = json.loads(text)
        ok = total = 0
        images = []
        for color, n in ((0, 1), (2, 3), (4, 2), (6, 4)):
            for depth in (8, 16):
                s = rng.integers(0, 1 << depth, (5, 7, n))
                images.append(("png", png_bytes(s, color, depth, interlace=bool(depth == 8))))
        for depth in (1, 2, 4, 8):
            s = rng.integers(0, 1 << depth, (6, 9, 1))
            images.append(("png", png_bytes(s, 0, depth)))
            pal = rng.integers(0, 256, (1 << depth, 3))
            images.append(("png", png_bytes(s, 3, depth, plte=pal, trns=bytes([7, 200]))))
        for rle in (False, True):
            images.append(("hdr", rgbe_bytes(rng.integers(0, 256, (4, 40, 4)), rle)))
        for kind, data in images:
            for k in range(1 + n_mut // 40):
                blob = data if k == 0 else mutate(rng, data, int(rng.integers(1, 4)))
                name = f"sky.{kind}"
                with open(os.path.join(tmp, name), "wb") as f:
                    f.write(blob)
                doc["GLOBALS"]["S"]["Skybox"] = name
                with open(path, "w") as f:
                    json.dump(doc, f)
                total += 1
                ok += load(path) == 0
        print(f"images: {total} sky images ({len(images)} valid kinds), {ok} accepted", flush=True)
    # procedural scenes, the sky image setter, the .hdr writer
    for mk in (lambda h: HOST.pico_scene_cornell(1, 1, h), lambda h: HOST.pico_scene_shapes(2, 0, h),
               lambda h: HOST.pico_scene_instanced(3, 4, 32, 100, 800, 10.0, 1, h),
               lambda h: HOST.pico_scene_training(10001, 1, h)):
        h = vp()
        assert mk(C.byref(h)) == 0
        d = pa.SceneDesc()
        assert HOST.pico_scene_desc_get(h, C.byref(d)) == 0
        tex = np.zeros((6, 3, 5, 4), np.uint8)
        img = pa.SkyImage(5, 3, 6, pa.TEXEL_RGBA8, tex.ctypes.data)
        assert HOST.pico_scene_set_sky_image(h, C.byref(img)) == 0
        HOST.pico_scene_destroy(h)
    with tempfile.TemporaryDirectory() as tmp:
        px = rng.gamma(1.0, 1.0, (9 * 7, 3)).astype(np.float32)
        assert HOST.pico_write_hdr(os.fsencode(os.path.join(tmp, "a.hdr")), 9, 7, px.ctypes.data, 1) == 0
    print("scenes: procedural scenes, sky images and the .hdr writer", flush=True)
    oracle_cases()
    print("done", flush=True)


def oracle_cases():
    """The oracle's scene build, closest hits, render, G-buffer, filters and
    CNN on the sanitized library."""
    import oracle
    oracle.CPU_REF_PATH = os.path.join(ROOT, "oracle", "_san", "libcpu_ref_san.so")
    h = vp()
    assert HOST.pico_scene_cornell(1, 1, C.byref(h)) == 0

    class S:  # what oracle.CpuScene needs of a scene
        pass
    sc = S()
    sc.desc = pa.SceneDesc()
    HOST.pico_scene_desc_get(h, C.byref(sc.desc))
    cs = oracle.CpuScene(sc)
    cam = pa.Camera()
    HOST.pico_scene_default_camera(h, 1.0, C.byref(cam))
    rng = np.random.default_rng(3)
    o = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (200, 1))
    d = rng.normal(size=(200, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cs.closest_hit(o, d)
    cs.closest_hit(o, d, brute=True)
    cs.primary_ids(cam, 16, 16)
    p = pa.params(12, 10, 3, 5, frame_seed=2)
    for rng_kind in (oracle.RNG_PCG, oracle.RNG_REFERENCE):
        cs.render(cam, p, rng=rng_kind, threads=2)
    n, pos, alb = cs.gbuffer(cam, 12, 10)
    img = rng.gamma(1.0, 1.0, (120, 3)).astype(np.float32)
    oracle.atrous(img, n, pos, alb, 12, 10, (4, 4), 5)
    oracle.atrous_fixed(img, n, pos, alb, 12, 10, 5)
    oracle.tonemap(img, 12, 10)
    par = rng.normal(0, 0.05, oracle.cnn_num_params())
    x = rng.uniform(0, 1, (10, 8, 8)).astype(np.float32)
    oracle.cnn_forward(par, x, 8, 8)
    oracle.cnn_forward(par, x[:, :7, :6].copy(), 7, 6, bf16=True)
    oracle.cnn_loss_grad(par, x[None], rng.uniform(0, 1, (1, 64, 3)).astype(np.float32), 1, 8, 8, bf16=True)
    del cs
    HOST.pico_scene_destroy(h)
    print("oracle: scene build, closest hits, render (both generators), G-buffer, filters, CNN", flush=True)


if __name__ == "__main__":
    main()
