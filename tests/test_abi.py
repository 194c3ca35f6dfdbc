"""CPU-side checks: the C-ABI library loads and exports every symbol the
header declares; the oracle libraries load; procedural scenes are well formed."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import pico_amd as pa


def test_library_exports_every_declared_symbol():
    lib = pa.lib()
    names = pa.exported_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_has_no_torch_types():
    hdr = open(os.path.join(os.path.dirname(pa.__file__), "..", "include", "pico_hip.h")).read()
    assert "torch" not in hdr and "at::" not in hdr


def test_oracle_libraries_load():
    assert oracle.cpu_lib().cpuref_pixel_seed(1, 2) != oracle.cpu_lib().cpuref_pixel_seed(1, 3)


def _arrays(scene):
    d = scene.desc
    u = lambda p, n: np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), (n,))
    f = lambda p, n: np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), (n,))
    vofs = u(d.mesh_vertex_offset, d.n_meshes + 1)
    tofs = u(d.mesh_triangle_offset, d.n_meshes + 1)
    return d, vofs, tofs, f(d.positions, 3 * vofs[-1]).reshape(-1, 3), f(d.normals, 3 * vofs[-1]).reshape(-1, 3), \
        u(d.indices, 3 * tofs[-1]).reshape(-1, 3)


def test_instanced_scene_matches_configuration():
    s = pa.Scene.instanced()
    d, vofs, tofs, pos, nrm, idx = _arrays(s)
    assert d.n_instances == 1024 + 3 and d.n_lights == 3
    per = np.diff(tofs)[:16]
    assert 1900 <= per.min() and per.max() <= 21000
    unique, instanced = s.n_triangles
    assert 7.5e6 < instanced < 9.5e6
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    for m in range(16):
        tri = idx[tofs[m]:tofs[m + 1]]
        assert tri.max() < vofs[m + 1] - vofs[m]
        # closed surface: every undirected edge belongs to exactly two triangles
        e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
        _, counts = np.unique(e, axis=0, return_counts=True)
        assert (counts == 2).all(), m


def test_scenes_are_deterministic():
    a = pa.Scene.instanced(5, 4, 10, 100, 400, 10.0)
    b = pa.Scene.instanced(5, 4, 10, 100, 400, 10.0)
    for x, y in zip(_arrays(a)[3:], _arrays(b)[3:]):
        assert np.array_equal(x, y)


def test_cornell_layout():
    s = pa.Scene.cornell(1)
    d, vofs, tofs, pos, nrm, idx = _arrays(s)
    assert d.n_instances == 8 and d.n_lights == 1
    bsdf = np.ctypeslib.as_array(C.cast(d.instance_bsdf, C.POINTER(C.c_uint32)), (8,))
    assert list(bsdf) == [pa.BSDF_DIELECTRIC] * 7 + [pa.BSDF_LIGHT]
    mats = np.ctypeslib.as_array(C.cast(d.materials, C.POINTER(C.c_float)), (d.n_materials * 11,)).reshape(-1, 11)
    assert C.sizeof(pa.Material) == 44 and (mats[:, 10] == 0).all()  # no index of refraction: opaque
    alb = mats[:7, 0:3] / 0.96
    assert (alb >= 0.2 - 1e-6).all() and (alb <= 0.8 + 1e-6).all()
    assert ((mats[:7, 6] >= 0.1) & (mats[:7, 6] <= 0.9)).all()
    np.testing.assert_array_equal(mats[7, 7:10], [10, 10, 10])


def test_gpu_entry_points_fail_cleanly_without_device():
    # On a machine without a GPU creating a context fails with an error code,
    # it does not crash; on a GPU machine it succeeds.
    h = C.c_void_p()
    rc = pa.lib().pico_hip_create(0, C.byref(h))
    if rc == 0:
        pa.lib().pico_hip_destroy(h)
    else:
        assert rc in (pa.PICO_ERR_HIP, pa.PICO_ERR_ARG)


def test_ctypes_layouts_match_the_header(tmp_path):
    """Every struct the binding declares has the C header's size and field
    offsets (compiled with the host C compiler)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"pico_scene_desc": pa.SceneDesc, "pico_sky_image": pa.SkyImage, "pico_material": pa.Material,
               "pico_camera": pa.Camera}
    inc = os.path.join(os.path.dirname(pa.__file__), "..", "include")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pico_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", inc, str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(py, f[0]).offset, (cname, f[0])
