"""The C++ front (pico_amd/front/pico_front.hpp) compiles against the C ABI,
links against libpico_hip.so, loads a JSON scene file with its named camera
(SceneFile, as Scene::Scene(ThreadPool&, path) and Scene::get_camera) and,
with a GPU, renders a denoised, tone-mapped Cornell frame and the scene file
with the call shape of Scene::render_scene_to_memory."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "pico_amd")


def _build(tmp_path):
    exe = str(tmp_path / "front_check")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", os.path.join(LIBDIR, "front", "front_check.cpp"), "-o", exe,
           f"-L{LIBDIR}", "-lpico_hip", f"-Wl,-rpath,{LIBDIR}", "-pthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_front_compiles_and_runs(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith(("ok", "nodevice")), r.stdout


@pytest.mark.gpu
def test_front_renders_on_device(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
