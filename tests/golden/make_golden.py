"""Writes tests/golden/ref_vectors.npz: input/output vectors of Pico's own
hot-path functions, computed by the reference subset compiled from the Pico
source tree (oracle/_ref/libpico_ref.so, built by oracle/Makefile where the
reference tree is present).  The fixtures let the oracle and the device path
be checked against the reference where the compiled subset is absent.

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
import pico_amd as pa  # noqa: E402
from sky_scenes import shapes_sun_sky, sun_cube_sky, sun_hdri_sky  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")


def scenes():
    return {
        "cornell_beckmann": pa.Scene.cornell(1, pa.MICROFACET_BECKMANN_ALL),
        "instanced_small": pa.Scene.instanced(7, 4, 48, 200, 2000, 12.0, pa.MICROFACET_BECKMANN_ALL),
        "cornell_glass": pa.Scene.cornell_glass(1, pa.MICROFACET_BECKMANN_ALL),
        "shapes": pa.Scene.shapes(1, pa.MICROFACET_BECKMANN_ALL),
        "sun_cube_sky": sun_cube_sky(),
        "sun_hdri_sky": sun_hdri_sky(),
        "shapes_sun_sky": shapes_sun_sky(),
    }


def scene_hash(s):
    d = s.desc
    nv = np.ctypeslib.as_array(oracle.C.cast(d.mesh_vertex_offset, oracle.C.POINTER(oracle.C.c_uint32)),
                               (d.n_meshes + 1,))[-1]
    pos = np.ctypeslib.as_array(oracle.C.cast(d.positions, oracle.C.POINTER(oracle.C.c_float)), (3 * int(nv),))
    xf = np.ctypeslib.as_array(oracle.C.cast(d.transforms, oracle.C.POINTER(oracle.C.c_float)), (16 * d.n_instances,))
    extra = b""
    if d.use_sun:
        extra += bytes(d.sun_direction) + bytes(d.sun_colour)
    if d.sky_image:
        k = pa.SkyImage.from_address(d.sky_image)
        n = k.width * k.height * k.faces * (16 if k.format == pa.TEXEL_RGBA32F else 4)
        extra += bytes(k)[:16] + oracle.C.string_at(k.texels, n)
    return hashlib.sha256(pos.tobytes() + xf.tobytes() + extra).hexdigest()[:16]


def main():
    if not oracle.ref_available():
        sys.exit("compiled reference subset not built (oracle/Makefile)")
    R = oracle.ref_lib()
    P = oracle._p
    g = {}
    # generators
    for seed in [0, 1, 12345, 0xdeadbeef]:
        b = np.empty(256, np.uint32)
        R.ref_xorshift(seed, 256, P(b))
        g[f"xorshift_{seed}"] = b
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2**32, size=256, dtype=np.uint64).astype(np.uint32)
    bits[:4] = [0, 1, 0xffffffff, 0x80000000]
    hb = np.empty((256, 2), np.float32)
    R.ref_hammersley(256, P(bits), P(hb))
    g["hammersley_in"], g["hammersley_out"] = bits, hb
    # distributions
    for kind, name in [(oracle.DIST_COS, "cos"), (oracle.DIST_BECKMANN, "beckmann")]:
        rng = np.random.default_rng(kind)
        n = 512
        xi = rng.uniform(0, 1, size=(n, 2)).astype(np.float32)
        v = rng.normal(size=(n, 3)).astype(np.float32)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        r = rng.uniform(0.05, 1.0, size=n).astype(np.float32)
        h, pdf = np.empty((n, 3), np.float32), np.empty(n, np.float32)
        R.ref_dist_sample(kind, n, P(xi), P(v), P(r), P(h))
        R.ref_dist_pdf(kind, n, P(v), P(h), P(r), P(pdf))
        g.update({f"dist_{name}_xi": xi, f"dist_{name}_v": v, f"dist_{name}_r": r, f"dist_{name}_h": h,
                  f"dist_{name}_pdf": pdf})
    # tangent frame
    rng = np.random.default_rng(3)
    n = 256
    v = rng.normal(size=(n, 3)).astype(np.float32)
    nn = rng.normal(size=(n, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    v[:16] = nn[:16]
    v[16:32] = -nn[16:32]
    fr = np.empty((n, 9), np.float32)
    R.ref_tangent_frame(n, P(v), P(nn), P(fr))
    g.update({"frame_v": v, "frame_n": nn, "frame_out": fr})
    # alias table
    rng = np.random.default_rng(4)
    for k in [1, 2, 5, 64]:
        w = rng.uniform(0.01, 1.0, size=k).astype(np.float32)
        s = np.empty(512, np.uint32)
        R.ref_alias_sample(k, P(w), 99, 512, P(s))
        g[f"alias_{k}_w"], g[f"alias_{k}_samples"] = w, s
    # BSDF sampling with the reference's generators
    m = pa.Material((0.4, 0.5, 0.6), (0.04, 0.04, 0.04), 0.35, (2.0, 2.0, 2.0))
    for bsdf in [pa.BSDF_DIFFUSE, pa.BSDF_SPECULAR, pa.BSDF_DIELECTRIC, pa.BSDF_LIGHT, pa.BSDF_DELTA]:
        rng = np.random.default_rng(10 + bsdf)
        n = 256
        nrm = rng.normal(size=(n, 3)).astype(np.float32)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        v = nrm + 0.8 * rng.normal(size=(n, 3)).astype(np.float32)
        v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
        l, p, e = np.empty((n, 3), np.float32), np.empty(n, np.float32), np.empty((n, 3), np.float32)
        R.ref_bsdf_sample(bsdf, m, n, P(v), P(nrm), 777, P(l), P(p), P(e))
        g.update({f"bsdf_{bsdf}_v": v, f"bsdf_{bsdf}_n": nrm, f"bsdf_{bsdf}_l": l, f"bsdf_{bsdf}_p": p,
                  f"bsdf_{bsdf}_e": e})
    # transmissive BSDFs and the IOR stack (entering from either side, leaving, nested)
    mg = pa.Material((0.9, 0.95, 1.0), (0.5, 0.5, 0.5), 0.3, (0, 0, 0), 1.45)
    for bsdf in [pa.BSDF_TRANSPARENT, pa.BSDF_FRESNEL]:
        rng = np.random.default_rng(30 + bsdf)
        n = 512
        nrm = rng.normal(size=(n, 3)).astype(np.float32)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        side = np.where(rng.uniform(size=n) < 0.5, 1.0, -1.0).astype(np.float32)
        v = side[:, None] * nrm + 0.9 * rng.normal(size=(n, 3)).astype(np.float32)
        v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
        inside = (((v * nrm).sum(1) < 0) | (rng.uniform(size=n) < 0.25)).astype(np.uint8)
        l, p, e = np.empty((n, 3), np.float32), np.empty(n, np.float32), np.empty((n, 3), np.float32)
        after = np.empty(n, np.uint8)
        R.ref_btdf_sample(bsdf, mg, n, P(v), P(nrm), P(inside), 91, P(l), P(p), P(e), P(after))
        g.update({f"btdf_{bsdf}_v": v, f"btdf_{bsdf}_n": nrm, f"btdf_{bsdf}_inside": inside, f"btdf_{bsdf}_l": l,
                  f"btdf_{bsdf}_p": p, f"btdf_{bsdf}_e": e, f"btdf_{bsdf}_after": after})
    # camera
    cam = pa.Camera((1.0, 2.0, -3.0), (0.2, -0.1, 1.0), (0.0, 1.0, 0.0), 1.5, 1000.0)
    rng = np.random.default_rng(2)
    n = 256
    px = rng.integers(0, 320, n).astype(np.uint32)
    py = rng.integers(0, 200, n).astype(np.uint32)
    xi = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    d = np.empty((n, 3), np.float32)
    R.ref_camera_rays(cam, 320, 200, n, P(px), P(py), P(xi), P(d))
    g.update({"camera_px": px, "camera_py": py, "camera_xi": xi, "camera_dir": d})
    # scene queries, integrator, accumulation, G-buffer
    for name, s in scenes().items():
        rs = oracle.RefScene(s)
        g[f"{name}_hash"] = np.frombuffer(scene_hash(s).encode(), np.uint8)
        cam = s.camera(1.0)
        inst, prim = rs.primary_ids(cam, 48, 48, (0.25, 0.75))
        g[f"{name}_ids_inst"], g[f"{name}_ids_prim"] = inst, prim
        rng = np.random.default_rng(1)
        n = 512
        o = rng.uniform(-1, 13, (n, 3)).astype(np.float32) if name in ("instanced_small", "sun_cube_sky") else \
            rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
        dd = rng.normal(size=(n, 3)).astype(np.float32)
        dd /= np.linalg.norm(dd, axis=1, keepdims=True)
        t, hi, hp, pos, nrm = rs.closest_hit(o, dd)
        g.update({f"{name}_ray_o": o, f"{name}_ray_d": dd, f"{name}_hit_t": t, f"{name}_hit_inst": hi,
                  f"{name}_hit_prim": hp, f"{name}_hit_pos": pos, f"{name}_hit_nrm": nrm})
        p = pa.params(16, 16, 6, 5, frame_seed=21)
        g[f"{name}_samples"] = rs.render_samples(cam, p)
        nb, pb, ab = rs.gbuffer(cam, 32, 24)
        g[f"{name}_gb_n"], g[f"{name}_gb_p"], g[f"{name}_gb_a"] = nb, pb, ab
        if name == "cornell_beckmann":
            p = pa.params(16, 16, 12, 4, frame_seed=5)
            pix, cnt, var = rs.render(cam, p, threads=1)
            g.update({"accum_pix": pix, "accum_cnt": cnt, "accum_var": var})
        del rs
    # sky lookup (ImageCube::sample4): a cube of RGBA8 faces, an equirectangular float image
    from sky_scenes import sky_dirs
    d = sky_dirs(2048, 7)
    rng = np.random.default_rng(12)
    for name, tex in [("cube", rng.integers(0, 256, (6, 11, 9, 4), dtype=np.uint8)),
                      ("hdri", rng.random((1, 13, 27, 4), dtype=np.float32))]:
        img = pa.SkyImage(tex.shape[2], tex.shape[1], tex.shape[0],
                          pa.TEXEL_RGBA32F if tex.dtype == np.float32 else pa.TEXEL_RGBA8, tex.ctypes.data)
        sky = (oracle.C.c_float * 3)(0.0, 0.0, 0.0)
        out = np.empty((len(d), 3), np.float32)
        R.ref_sky_sample(oracle.C.addressof(img), oracle.C.addressof(sky), len(d), P(d), P(out))
        g.update({f"sky_{name}_tex": tex, f"sky_{name}_dir": d, f"sky_{name}_out": out})
    # a-trous and tone mapping
    rng = np.random.default_rng(8)
    w, h = 41, 31
    rgb = rng.gamma(2.0, 0.5, (w * h, 3)).astype(np.float32)
    n = rng.normal(size=(w * h, 3)).astype(np.float32)
    pos = (rng.normal(size=(w * h, 3)) * 4).astype(np.float32)
    alb = rng.uniform(0, 1, (w * h, 3)).astype(np.float32)
    g.update({"atrous_rgb": rgb, "atrous_n": n, "atrous_p": pos, "atrous_a": alb})
    for tile in [(16, 16), (5, 7)]:
        g[f"atrous_out_{tile[0]}x{tile[1]}"] = oracle.ref_atrous(rgb, n, pos, alb, w, h, tile)
    rng = np.random.default_rng(9)
    tm = rng.gamma(1.5, 2.0, (64 * 32, 3)).astype(np.float32)
    g["tonemap_in"], g["tonemap_out"] = tm, oracle.ref_tonemap(tm, 64, 32)
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT}: {len(g)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
