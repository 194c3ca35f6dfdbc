"""Model files through the reference's import path (Scene::Scene(ThreadPool&,
dir, aiScene*, options) with parse_node and add_material,
Source/Core/Scene.cpp:86-141, 722-971): an OBJ file split into meshes as
assimp's OBJ importer splits it, every mesh an identity-transform instance,
Light when its material emits and Dielectric otherwise, materials from the
mtllib as the constant branch of add_material reads them.  The reference's
loader cannot run here (assimp needs CMake and generated headers), so the
import rules are restated and checked case by case; the geometry is checked
against brute-force closest hits.  Host only."""
import ctypes as C

import numpy as np
import pytest

import oracle
import pico_amd as pa

MTL = """# materials
newmtl Grey
Kd 0.5 0.5 0.5
Ks 0.04 0.04 0.04
newmtl Red
Kd 0.7 0.1 0.1
newmtl Blue
Kd 0.1 0.1 0.7
Ks 0.2
newmtl Lamp
Kd 0.8 0.8 0.8
Ke 6 6 5
"""

# a Cornell-like room: floor, back wall, a box whose top is blue, a lamp
ROOM = """mtllib room.mtl
v -1 -1 -1
v 1 -1 -1
v 1 -1 1
v -1 -1 1
v -1 1 1
v 1 1 1
o room
usemtl Grey
f 1 4 3 2
f 4 5 6 3
o box
usemtl Red
v -0.4 -1 -0.2
v 0.2 -1 -0.2
v 0.2 -1 0.4
v -0.4 -1 0.4
v -0.4 -0.3 -0.2
v 0.2 -0.3 -0.2
v 0.2 -0.3 0.4
v -0.4 -0.3 0.4
f 7 8 12 11
f 8 9 13 12
f 9 10 14 13
f 10 7 11 14
usemtl Blue
f -4 -3 -2 -1
o lamp
usemtl Lamp
v -0.25 0.99 -0.25
v 0.25 0.99 -0.25
v 0.25 0.99 0.25
v -0.25 0.99 0.25
f -4 -1 -2 -3
"""


def room(tmp_path, text=ROOM, mtl=MTL):
    (tmp_path / "room.mtl").write_text(mtl)
    p = tmp_path / "room.obj"
    p.write_text(text)
    return p


def _desc(scene):
    d = scene.desc
    u32 = C.POINTER(C.c_uint32)
    mats = [(list(m.diffuse), list(m.specular), m.roughness, list(m.emissive))
            for m in (pa.Material * d.n_materials).from_address(d.materials)]
    return {"tofs": np.ctypeslib.as_array(C.cast(d.mesh_triangle_offset, u32), (d.n_meshes + 1,)).copy(),
            "imesh": np.ctypeslib.as_array(C.cast(d.instance_mesh, u32), (d.n_instances,)).copy(),
            "imat": np.ctypeslib.as_array(C.cast(d.instance_material, u32), (d.n_instances,)).copy(),
            "bsdf": np.ctypeslib.as_array(C.cast(d.instance_bsdf, u32), (d.n_instances,)).copy(),
            "xf": np.ctypeslib.as_array(C.cast(d.transforms, C.POINTER(C.c_float)), (16 * d.n_instances,)).copy(),
            "lights": np.ctypeslib.as_array(C.cast(d.light_instances, u32), (d.n_lights,)).copy()
            if d.n_lights else np.zeros(0, np.uint32),
            "mats": mats, "microfacet": d.microfacet, "sun": d.use_sun, "sky": list(d.sky_rgb)}


def test_meshes_materials_and_lights(tmp_path):
    d = _desc(pa.Scene.load_obj(room(tmp_path)))
    # room (grey), box sides (red), box top (blue, a material change), lamp
    assert np.diff(d["tofs"]).tolist() == [4, 8, 2, 2]
    assert d["bsdf"].tolist() == [pa.BSDF_DIELECTRIC] * 3 + [pa.BSDF_LIGHT]
    assert d["lights"].tolist() == [3]
    assert d["imesh"].tolist() == [0, 1, 2, 3]
    assert np.array_equal(d["xf"].reshape(-1, 4, 4), np.tile(np.eye(4, dtype=np.float32), (4, 1, 1)))
    assert d["microfacet"] == pa.MICROFACET_BECKMANN_ALL and not d["sun"] and d["sky"] == [0, 0, 0]
    m = [d["mats"][i] for i in d["imat"]]
    assert m[0][0] == pytest.approx([0.5] * 3) and m[0][1] == pytest.approx([0.04] * 3) and m[0][3] == [0, 0, 0]
    assert m[1][0] == pytest.approx([0.7, 0.1, 0.1]) and m[1][1] == [0, 0, 0]
    assert m[2][0] == pytest.approx([0.1, 0.1, 0.7]) and m[2][1] == pytest.approx([0.2] * 3)  # one value: all three
    assert m[3][3] == [6, 6, 5] and all(x[2] == 0.0 for x in m)  # roughness 1 - 1^2


def test_splitting_rules(tmp_path):
    text = """mtllib room.mtl
v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
f 1 2 3
usemtl Red
usemtl Red
f 1 2 4
g a
f 1 3 4
g a
f 2 3 4
o named
usemtl Blue
usemtl Grey
f 1 2 3
o other
f 1 2 4
o named
f 1 3 4
usemtl Nowhere
f 2 3 4
"""
    d = _desc(pa.Scene.load_obj(room(tmp_path, text)))
    # default object: (default material, 1 face), Red (1); group a: 2 faces in
    # one mesh; object named: Grey (2 faces, continued after "other"),
    # Nowhere (1); other: Grey (1)
    assert np.diff(d["tofs"]).tolist() == [1, 1, 2, 2, 1, 1]
    mats = [d["mats"][i] for i in d["imat"]]
    assert mats[0][0] == pytest.approx([0.6] * 3)          # assimp's default material
    assert mats[1][0] == pytest.approx([0.7, 0.1, 0.1])    # Red
    assert mats[3][0] == pytest.approx([0.5] * 3)          # named: Grey
    assert mats[4][0] == pytest.approx([0.6] * 3)          # a material no library defines
    assert mats[5][0] == pytest.approx([0.5] * 3)          # other: Grey carried over
    assert d["lights"].size == 0


def test_room_hits_match_brute_force(tmp_path):
    s = pa.Scene.load_obj(room(tmp_path))
    rng = np.random.default_rng(3)
    o = rng.uniform(-0.9, 0.9, (3000, 3)).astype(np.float32)
    v = rng.normal(size=(3000, 3))
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    cs = oracle.CpuScene(s)
    t, i, p, *_ = cs.closest_hit(o, v)
    bt, bi, bp, *_ = cs.closest_hit(o, v, brute=True)
    assert (bi != 0xffffffff).mean() > 0.3  # floor, back wall and box; the rest is open
    assert ((i == bi) & (p == bp)).mean() > 0.999


@pytest.mark.parametrize("text,needle", [("v 0 0 0\n", "no faces"), ("v 0 0 0\nf 1 2 3\n", "out of range"),
                                         ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 x 3\n", "malformed face")])
def test_errors(tmp_path, text, needle):
    with pytest.raises(pa.PicoError) as e:
        pa.Scene.load_obj(room(tmp_path, text))
    assert needle in str(e.value)


def test_missing_library_and_file(tmp_path):
    with pytest.raises(pa.PicoError):
        pa.Scene.load_obj(tmp_path / "none.obj")
    p = tmp_path / "m.obj"
    p.write_text("mtllib nothere.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(pa.PicoError) as e:
        pa.Scene.load_obj(p)
    assert "nothere.mtl" in str(e.value)
