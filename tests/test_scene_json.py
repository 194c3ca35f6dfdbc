"""Scene files in the reference's JSON format (Source/Core/Scene.cpp:36-84,
297-720) read into the flattened scene: OBJ meshes (triangulation, flat
normals, joined corners, first mesh only), instance transforms T * R * S,
the two Constant material forms, BSRDF kinds, light and camera order, path
resolution, and the errors for what is not supported.  Host only."""
import ctypes as C
import json
import math

import numpy as np
import pytest

import oracle
import pico_amd as pa


def _desc(scene):
    d = scene.desc
    nm, ni = d.n_meshes, d.n_instances
    u32 = lambda p, n: np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), (n,)).copy() if n else np.zeros(0, np.uint32)
    f32 = lambda p, n: np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), (n,)).copy() if n else np.zeros(0, np.float32)
    vofs, tofs = u32(d.mesh_vertex_offset, nm + 1), u32(d.mesh_triangle_offset, nm + 1)
    mats = [(tuple(m.diffuse), tuple(m.specular), m.roughness, tuple(m.emissive))
            for m in (pa.Material * d.n_materials).from_address(d.materials)]
    ior = [m.index_of_refraction for m in (pa.Material * d.n_materials).from_address(d.materials)]
    return {
        "vofs": vofs, "tofs": tofs,
        "pos": f32(d.positions, 3 * int(vofs[-1])).reshape(-1, 3),
        "nrm": f32(d.normals, 3 * int(vofs[-1])).reshape(-1, 3),
        "idx": u32(d.indices, 3 * int(tofs[-1])).reshape(-1, 3),
        "xf": f32(d.transforms, 16 * ni).reshape(ni, 4, 4).transpose(0, 2, 1),  # row-major matrices
        "mesh": u32(d.instance_mesh, ni), "mat": u32(d.instance_material, ni), "bsdf": u32(d.instance_bsdf, ni),
        "lights": u32(d.light_instances, d.n_lights), "mats": mats, "ior": ior, "sky": tuple(d.sky_rgb), "microfacet": d.microfacet,
    }


def _write(tmp_path, doc, files):
    for name, text in files.items():
        p = tmp_path / name
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(text)
    path = tmp_path / "scene.json"
    path.write_text(doc if isinstance(doc, str) else json.dumps(doc, indent=1))
    return path


QUAD = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nf 1 2 3 4\n"
MAT = {"Grey": {"Type": "Constant", "Albedo": [0.5, 0.5, 0.5], "Metalness": 0.0, "Roughness": 0.5}}


def _scene(meshes, instances, materials=MAT, cameras=None, globals_=None):
    doc = {"MESH": {k: {"Path": v} for k, v in meshes.items()}, "MATERIALS": materials, "INSTANCE": instances}
    if cameras is not None:
        doc["CAMERA"] = cameras
    if globals_ is not None:
        doc["GLOBALS"] = globals_
    return doc


def test_quad_is_a_fan_with_flat_normals(tmp_path):
    path = _write(tmp_path, _scene({"Quad": "quad.obj"}, {"A": {"Asset": "Quad", "BSRDF": "Diffuse"}}), {"quad.obj": QUAD})
    d = _desc(pa.Scene.load_json(path))
    # the polygon becomes (0, 1, 2), (0, 2, 3); the flat normal of both is
    # normalize((v1 - v0) x (v2 - v0)) = (0, -1, 0), so the shared corners join
    np.testing.assert_array_equal(d["idx"], [[0, 1, 2], [0, 2, 3]])
    np.testing.assert_array_equal(d["pos"], [[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]])
    np.testing.assert_array_equal(d["nrm"], np.tile([0.0, -1.0, 0.0], (4, 1)))
    assert d["microfacet"] == pa.MICROFACET_BECKMANN_ALL
    assert d["sky"] == (0.0, 0.0, 0.0)


def test_obj_normals_negative_indices_and_first_mesh_only(tmp_path):
    obj = ("# a triangle with normals, given by relative indices\n"
           "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvt 0 0\n"
           "f -3/1/-1 -2/1/-1 -1/1/-1\n"
           "o second\nv 5 5 5\nv 6 5 5\nv 5 6 5\nf 4 5 6\n")
    path = _write(tmp_path, _scene({"T": "t.obj"}, {"A": {"Asset": "T", "BSRDF": "Diffuse"}}), {"t.obj": obj})
    d = _desc(pa.Scene.load_json(path))
    assert d["idx"].shape == (1, 3)
    np.testing.assert_array_equal(d["pos"], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    np.testing.assert_array_equal(d["nrm"], np.tile([0.0, 0.0, 1.0], (3, 1)))


def test_corners_split_by_normal(tmp_path):
    # two faces of a cube sharing an edge, each with its own normal: the
    # shared positions appear once per normal
    obj = ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 1 0 1\nv 1 1 1\nvn 0 0 -1\nvn 1 0 0\n"
           "f 1//1 4//1 3//1 2//1\nf 2//2 3//2 6//2 5//2\n")
    path = _write(tmp_path, _scene({"C": "c.obj"}, {"A": {"Asset": "C", "BSRDF": "Diffuse"}}), {"c.obj": obj})
    d = _desc(pa.Scene.load_json(path))
    assert len(d["pos"]) == 8 and d["idx"].shape == (4, 3)
    # every triangle's corners carry its face's normal
    for tri in d["idx"]:
        n = d["nrm"][tri]
        assert (n == n[0]).all()


def _trs(p, q, s):
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(q)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(s)
    m[:3, 3] = p
    return m


def test_instance_transform_is_translate_rotate_scale(tmp_path):
    q = [0.0, 2 * math.sin(math.pi / 4), 0.0, 2 * math.cos(math.pi / 4)]  # 90 deg about +y, not normalised
    inst = {"A": {"Asset": "Quad", "BSRDF": "Diffuse", "Position": [1, 2, 3], "Scale": [2, 3, 4], "Rotation": q},
            "B": {"Asset": "Quad", "BSRDF": "Diffuse"}}
    path = _write(tmp_path, _scene({"Quad": "quad.obj"}, inst), {"quad.obj": QUAD})
    d = _desc(pa.Scene.load_json(path))
    np.testing.assert_allclose(d["xf"][0], _trs([1, 2, 3], q, [2, 3, 4]), atol=1e-6)
    # closed form: the scaled x axis turns onto -z, y stays, z turns onto +x
    np.testing.assert_allclose(d["xf"][0] @ [1, 0, 0, 1], [1, 2, 1, 1], atol=1e-6)
    np.testing.assert_allclose(d["xf"][0] @ [0, 1, 0, 1], [1, 5, 3, 1], atol=1e-6)
    np.testing.assert_allclose(d["xf"][0] @ [0, 0, 1, 1], [5, 2, 3, 1], atol=1e-6)
    np.testing.assert_array_equal(d["xf"][1], np.eye(4))  # defaults
    assert d["mesh"][0] == d["mesh"][1]  # one mesh, two instances


def test_materials(tmp_path):
    mats = {
        "Dielec": {"Type": "Constant", "Albedo": [0.2, 0.4, 0.8], "Metalness": 0.0, "Roughness": 0.3},
        "Metal": {"Type": "Constant", "Albedo": [0.2, 0.4, 0.8], "Metalness": 1.0, "Roughness": 0.1},
        "Glossy": {"Type": "Constant", "Diffuse": [0.1, 0.2, 0.3], "Specular": [0.5, 0.5, 0.5], "Gloss": 0.5},
        "Lamp": {"Type": "Constant", "Albedo": [0.8, 0.8, 0.8], "Roughness": 0.5, "Emissive": [10, 9, 8]},
        "Glass": {"Type": "Constant", "Albedo": [1, 1, 1], "Roughness": 0.0, "Emissive": [5, 5, 5],
                  "IndexOfRefraction": 1.5},
    }
    inst = {"I": {"Asset": "Quad", "BSRDF": "Dielectric", "Material": "Metal"}}
    path = _write(tmp_path, _scene({"Quad": "quad.obj"}, inst, materials=mats), {"quad.obj": QUAD})
    d = _desc(pa.Scene.load_json(path))
    m = dict(zip(sorted(mats), d["mats"]))
    f = np.float32
    # metalness 0: diffuse = 0.96 albedo, specular = 0.04
    np.testing.assert_allclose(m["Dielec"][0], f(0.96) * np.array([0.2, 0.4, 0.8], f), rtol=1e-6)
    np.testing.assert_allclose(m["Dielec"][1], [0.04] * 3, rtol=1e-6)
    assert m["Dielec"][2] == pytest.approx(0.3)
    # metalness 1: no diffuse, specular = albedo
    np.testing.assert_allclose(m["Metal"][0], [0, 0, 0], atol=1e-7)
    np.testing.assert_allclose(m["Metal"][1], [0.2, 0.4, 0.8], rtol=1e-6)
    # diffuse/specular/gloss: roughness = 1 - gloss^2
    assert m["Glossy"][0] == pytest.approx((0.1, 0.2, 0.3)) and m["Glossy"][1] == pytest.approx((0.5, 0.5, 0.5))
    assert m["Glossy"][2] == pytest.approx(0.75)
    assert m["Lamp"][3] == pytest.approx((10, 9, 8))
    assert m["Glass"][3] == (0.0, 0.0, 0.0)  # the transparent forms emit nothing
    assert d["ior"][sorted(mats).index("Glass")] == pytest.approx(1.5) and d["ior"][0] == 0.0
    assert d["mat"][0] == sorted(mats).index("Metal") and d["bsdf"][0] == pa.BSDF_DIELECTRIC


def test_order_lights_and_cameras(tmp_path):
    inst = {"c_light": {"Asset": "Quad", "BSRDF": "Light", "Material": "Lamp"},
            "a_floor": {"Asset": "Quad", "BSRDF": "Diffuse"},
            "b_light": {"Asset": "Quad", "BSRDF": "Light", "Material": "Lamp"},
            "d_mirror": {"Asset": "Quad", "BSRDF": "Delta"},
            "e_spec": {"Asset": "Quad", "BSRDF": "Specular"}}
    mats = dict(MAT, Lamp={"Type": "Constant", "Albedo": [1, 1, 1], "Emissive": [4, 4, 4]})
    cams = {"main": {"Position": [0, 1, -3], "Direction": [0, 0, 2], "Aspect": 1.5, "FarPlane": 50.0},
            "aux": {"Direction": [3, 0, 4]}}
    path = _write(tmp_path, _scene({"Quad": "quad.obj"}, inst, materials=mats, cameras=cams), {"quad.obj": QUAD})
    s = pa.Scene.load_json(path)
    d = _desc(s)
    # instances in name order; the second light on the mesh gets its own copy
    assert list(d["bsdf"]) == [pa.BSDF_DIFFUSE, pa.BSDF_LIGHT, pa.BSDF_LIGHT, pa.BSDF_DELTA, pa.BSDF_SPECULAR]
    assert list(d["lights"]) == [1, 2]
    assert d["mesh"][0] == d["mesh"][1] == d["mesh"][3] == d["mesh"][4] != d["mesh"][2]
    assert len(d["vofs"]) == 3
    cams_out = s.cameras()
    assert list(cams_out) == ["aux", "main"]
    aux, main = cams_out["aux"], cams_out["main"]
    assert tuple(aux.direction) == pytest.approx((0.6, 0.0, 0.8)) and tuple(aux.position) == (0, 0, 0)
    assert tuple(aux.up) == (0, 1, 0) and aux.aspect == 1.0 and aux.far_plane == 10.0  # Camera defaults
    assert tuple(main.direction) == (0, 0, 1) and main.aspect == 1.5 and main.far_plane == 50.0
    assert tuple(s.camera(2.0).direction) == pytest.approx((0.6, 0.0, 0.8))  # the first camera by name


def test_builtin_shapes(tmp_path):
    inst = {"a_ball": {"Asset": "Sphere", "BSRDF": "Diffuse", "Scale": [2, 2, 2]},
            "b_box": {"Asset": "Cube", "BSRDF": "Specular"},
            "c_ball": {"Asset": "Sphere", "BSRDF": "Dielectric"},
            "d_lamp": {"Asset": "Sphere", "BSRDF": "Light", "Material": "Lamp"},
            "e_lamp": {"Asset": "Sphere", "BSRDF": "Light", "Material": "Lamp"},
            "f_quad": {"Asset": "Quad", "BSRDF": "Diffuse"}}
    mats = dict(MAT, Lamp={"Type": "Constant", "Albedo": [1, 1, 1], "Emissive": [4, 4, 4]})
    path = _write(tmp_path, _scene({"Quad": "quad.obj"}, inst, materials=mats), {"quad.obj": QUAD})
    s = pa.Scene.load_json(path)
    d = _desc(s)
    shapes = np.ctypeslib.as_array(C.cast(s.desc.mesh_shape, C.POINTER(C.c_uint32)), (s.desc.n_meshes,))
    kinds = shapes[d["mesh"]]
    assert list(kinds) == [pa.SHAPE_SPHERE, pa.SHAPE_CUBE, pa.SHAPE_SPHERE, pa.SHAPE_SPHERE, pa.SHAPE_SPHERE,
                           pa.SHAPE_TRIANGLES]
    # one entry for the sphere; a second light on it gets its own (lights must not share one)
    assert d["mesh"][0] == d["mesh"][2] == d["mesh"][3] != d["mesh"][4]
    assert (np.diff(d["vofs"])[shapes != 0] == 0).all() and (np.diff(d["tofs"])[shapes != 0] == 0).all()
    # all at the origin: the ray along +z meets the unit cube's face at 4.5;
    # the sphere scaled by 2 is missed, as in the reference, whose quadratic
    # takes the object-space direction (here of length 0.5) as unit length
    t, i, *_ = oracle.CpuScene(s).closest_hit(np.array([[0, 0.1, -5]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert i[0] == 1 and abs(t[0] - 4.5) < 1e-5


def test_case_insensitive_paths(tmp_path):
    path = _write(tmp_path, _scene({"Quad": "Meshes\\QUAD.OBJ"}, {"A": {"Asset": "Quad", "BSRDF": "Diffuse"}}),
                  {"meshes/quad.obj": QUAD})
    assert _desc(pa.Scene.load_json(path))["idx"].shape == (2, 3)


def test_json_syntax(tmp_path):
    # escapes, exponents, a repeated member (the last one counts), whitespace
    text = ('{"MESH": {"Qu\\u0061d": {"Path": "quad.obj"}},\n "MATERIALS": {"M": {"Type": "Constant", '
            '"Albedo": [5e-1, 0.5E0, -0.0], "Roughness": 2.5e-1}},\n'
            '"INSTANCE": {"A": {"Asset": "Quad", "BSRDF": "Specular", "BSRDF": "Diffuse", "Position": [1e0, 0, 0]}}}')
    path = _write(tmp_path, text, {"quad.obj": QUAD})
    d = _desc(pa.Scene.load_json(path))
    assert d["bsdf"][0] == pa.BSDF_DIFFUSE and d["mats"][0][2] == 0.25 and d["xf"][0][0, 3] == 1.0
    # jsoncpp's extensions the reference's loader accepts: comments, trailing
    # commas, a byte-order mark, text after the document
    text2 = ("\ufeff// scene\n" + text[:-1] + ", /* more */ }  trailing").encode()
    (tmp_path / "scene.json").write_bytes(text2)
    d2 = _desc(pa.Scene.load_json(tmp_path / "scene.json"))
    assert d2["bsdf"][0] == pa.BSDF_DIFFUSE and d2["mats"] == d["mats"]


@pytest.mark.parametrize("case,needle", [
    ('{"MESH": {"Q": {"Path": "quad.obj"}},\n "MATERIALS": [1, 2 3]}', "scene.json:2:"),
    ("unknown_asset", "unknown asset"),
    ("sphere", "the name of a built-in shape"),
    ("transparent", "needs a material with an IndexOfRefraction"),
    ("image_material", "image materials"),
    ("skybox", "cannot read"),
    ("sun", "SunColour: expected an array of 3 numbers"),
    ("missing_file", "cannot open"),
    ("bad_face", "face index out of range"),
    ("no_bsdf", "BSRDF: expected a string"),
    ("unknown_material", "unknown material"),
])
def test_errors(tmp_path, case, needle):
    files = {"quad.obj": QUAD, "bad.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n"}
    inst = {"A": {"Asset": "Quad", "BSRDF": "Diffuse"}}
    meshes = {"Quad": "quad.obj"}
    mats, glob = MAT, None
    if case == "unknown_asset":
        inst = {"A": {"Asset": "Nope", "BSRDF": "Diffuse"}}
    elif case == "sphere":
        meshes = {"Sphere": "quad.obj"}
    elif case == "transparent":
        inst = {"A": {"Asset": "Quad", "BSRDF": "Transparent"}}
    elif case == "image_material":
        mats = {"Tex": {"Type": "Metalic", "Albedo": "albedo.png"}}
    elif case == "skybox":
        glob = {"Sky": {"Skybox": "sky.hdr"}}
    elif case == "sun":
        glob = {"Sky": {"SunDirection": [0, -1, 0], "SunColour": [1, 1]}}
    elif case == "missing_file":
        meshes = {"Quad": "nothere.obj"}
    elif case == "bad_face":
        meshes = {"Quad": "bad.obj"}
    elif case == "no_bsdf":
        inst = {"A": {"Asset": "Quad"}}
    elif case == "unknown_material":
        inst = {"A": {"Asset": "Quad", "BSRDF": "Diffuse", "Material": "Gold"}}
    doc = case if case.startswith("{") else _scene(meshes, inst, materials=mats, globals_=glob)
    path = _write(tmp_path, doc, files)
    with pytest.raises(pa.PicoError) as e:
        pa.Scene.load_json(path)
    assert e.value.code == pa.PICO_ERR_ARG and needle in str(e.value), str(e.value)


def cornell_json(tmp_path):
    """A Cornell box in the reference's scene format: walls turned by
    quaternions, two boxes from an OBJ with normals (one of Fresnel glass),
    one quad light."""
    cube = ["v -0.5 -0.5 -0.5", "v 0.5 -0.5 -0.5", "v 0.5 0.5 -0.5", "v -0.5 0.5 -0.5",
            "v -0.5 -0.5 0.5", "v 0.5 -0.5 0.5", "v 0.5 0.5 0.5", "v -0.5 0.5 0.5",
            "vn 0 0 -1", "vn 0 0 1", "vn -1 0 0", "vn 1 0 0", "vn 0 -1 0", "vn 0 1 0",
            "f 1//1 4//1 3//1 2//1", "f 5//2 6//2 7//2 8//2", "f 1//3 5//3 8//3 4//3",
            "f 2//4 3//4 7//4 6//4", "f 1//5 2//5 6//5 5//5", "f 4//6 8//6 7//6 3//6"]
    up = "v -1 0 -1\nv -1 0 1\nv 1 0 1\nv 1 0 -1\nf 1 2 3 4\n"  # faces +y
    s = math.sqrt(0.5)
    inst = {
        "floor": {"Asset": "Wall", "BSRDF": "Dielectric", "Material": "White", "Position": [0, -1, 0]},
        "ceiling": {"Asset": "Wall", "BSRDF": "Dielectric", "Material": "White", "Position": [0, 1, 0],
                    "Rotation": [1, 0, 0, 0]},
        "left": {"Asset": "Wall", "BSRDF": "Dielectric", "Material": "Red", "Position": [-1, 0, 0],
                 "Rotation": [0, 0, -s, s]},
        "right": {"Asset": "Wall", "BSRDF": "Dielectric", "Material": "Green", "Position": [1, 0, 0],
                  "Rotation": [0, 0, s, s]},
        "back": {"Asset": "Wall", "BSRDF": "Dielectric", "Material": "White", "Position": [0, 0, 1],
                 "Rotation": [-s, 0, 0, s]},
        "box1": {"Asset": "Box", "BSRDF": "Dielectric", "Material": "White", "Position": [-0.35, -0.7, 0.3],
                 "Rotation": [0, 0.2588, 0, 0.9659], "Scale": [0.6, 0.6, 0.6]},
        "box2": {"Asset": "Box", "BSRDF": "Fresnel", "Material": "Glass", "Position": [0.35, -0.4, -0.2],
                 "Rotation": [0, -0.2588, 0, 0.9659], "Scale": [0.6, 1.2, 0.6]},
        "light": {"Asset": "Lamp", "BSRDF": "Light", "Material": "Lamp", "Position": [0, 0.99, 0],
                  "Scale": [0.5, 1, 0.5], "Rotation": [1, 0, 0, 0]},
    }
    mats = {"White": {"Type": "Constant", "Albedo": [0.7, 0.7, 0.7], "Roughness": 0.6},
            "Red": {"Type": "Constant", "Albedo": [0.7, 0.1, 0.1], "Roughness": 0.5},
            "Green": {"Type": "Constant", "Albedo": [0.1, 0.7, 0.1], "Roughness": 0.4},
            "Metal": {"Type": "Constant", "Albedo": [0.9, 0.8, 0.6], "Metalness": 1.0, "Roughness": 0.3},
            "Glass": {"Type": "Constant", "Diffuse": [0.95, 0.95, 0.95], "Specular": [0.9, 0.9, 0.9], "Gloss": 0.95,
                      "IndexOfRefraction": 1.5},
            "Lamp": {"Type": "Constant", "Albedo": [0.8, 0.8, 0.8], "Roughness": 0.5, "Emissive": [12, 12, 12]}}
    cams = {"front": {"Position": [0, 0, -1], "Direction": [0, 0, 1], "Aspect": 1.0, "FarPlane": 100.0}}
    doc = {"GLOBALS": {"Settings": {}}, "MESH": {"Wall": {"Path": "wall.obj"}, "Box": {"Path": "box.obj"},
                                                "Lamp": {"Path": "lamp.obj"}},
           "MATERIALS": mats, "INSTANCE": inst, "CAMERA": cams}
    return _write(tmp_path, doc, {"wall.obj": up, "lamp.obj": up, "box.obj": "\n".join(cube) + "\n"})


def test_cornell_json_hits_match_brute_force(tmp_path):
    s = pa.Scene.load_json(cornell_json(tmp_path))
    d = _desc(s)
    assert len(d["bsdf"]) == 8 and list(d["lights"]) == [6]  # "light" is 7th in name order
    assert d["bsdf"][2] == pa.BSDF_FRESNEL  # box2 (back, box1, box2, ...)
    rng = np.random.default_rng(1)
    o = rng.uniform(-0.9, 0.9, (3000, 3)).astype(np.float32)
    v = rng.normal(size=(3000, 3))
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    cs = oracle.CpuScene(s)
    t, i, p, *_ = cs.closest_hit(o, v)
    bt, bi, bp, *_ = cs.closest_hit(o, v, brute=True)
    assert (bi != 0xffffffff).mean() > 0.8  # rays leave only through the open front
    assert ((i == bi) & (p == bp)).mean() > 0.999
