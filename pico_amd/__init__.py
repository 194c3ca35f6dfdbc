"""pico_amd — Python binding of the MI355X (gfx950) rendering hot path of Pico.

The compute path is the C ABI of ``libpico_hip.so`` (declared in
``include/pico_hip.h``); this module only wraps it with ctypes for tests,
benchmarks and scripts.  There is no fallback: if the library is missing, or
no GPU is visible, the calls raise.

Reference interfaces mirrored here (Pico source tree):
  * ``Context.render``   -> Scene::Scene::render_scene_to_memory (Source/Core/Scene.cpp:143-235)
  * ``Context.gbuffer``  -> Scene::generate_denoiser_inputs (Source/Core/Scene.cpp:248-294)
  * ``Context.atrous``   -> Util::atrous_denoise (Source/Util/Denoisers.cpp:107-126)
  * ``Context.tonemap``  -> Util::reinhard_tone_mapping (Source/Util/ToneMappers.cpp:12-44)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpico_hip.so")

PICO_OK = 0
BUILD_DEVICE = 1
PICO_ERR_ARG = -1
PICO_ERR_HIP = -2
PICO_ERR_STATE = -3
PICO_ERR_CANCELLED = -4
PICO_ERR_LIMIT = -5
PICO_ERR_NOMEM = -6

BSDF_DIFFUSE, BSDF_SPECULAR, BSDF_DIELECTRIC, BSDF_LIGHT, BSDF_DELTA, BSDF_TRANSPARENT, BSDF_FRESNEL = range(7)
MICROFACET_BECKMANN_ALL, MICROFACET_GGX = 0, 1
TEXEL_RGBA8, TEXEL_RGBA32F = 0, 1
SHAPE_TRIANGLES, SHAPE_SPHERE, SHAPE_CUBE = 0, 1, 2
RENDER_RESUME, RENDER_COUNT = 1, 2
TUNE_SIDE_STREAM = 1
TUNE_SHORTEST_FIRST = 2
# fixed-width a-trous: normal, relative depth, albedo, colour (0 = term off)
ATROUS_FIXED_SIGMAS = (0.5, 0.05, 0.1, 0.0)


class PicoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pico error {code}: {msg}")
        self.code = code


class Material(C.Structure):
    _fields_ = [("diffuse", C.c_float * 3), ("specular", C.c_float * 3), ("roughness", C.c_float),
                ("emissive", C.c_float * 3), ("index_of_refraction", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("n_meshes", C.c_uint32), ("mesh_vertex_offset", C.c_void_p), ("mesh_triangle_offset", C.c_void_p),
                ("positions", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
                ("n_instances", C.c_uint32), ("transforms", C.c_void_p), ("instance_mesh", C.c_void_p),
                ("instance_material", C.c_void_p), ("instance_bsdf", C.c_void_p), ("n_materials", C.c_uint32),
                ("materials", C.c_void_p), ("n_lights", C.c_uint32), ("light_instances", C.c_void_p),
                ("sky_rgb", C.c_float * 3), ("microfacet", C.c_uint32), ("mesh_shape", C.c_void_p),
                ("use_sun", C.c_uint32), ("sun_direction", C.c_float * 3), ("sun_colour", C.c_float * 3),
                ("sky_image", C.c_void_p)]


class SkyImage(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("faces", C.c_uint32), ("format", C.c_uint32),
                ("texels", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
                ("aspect", C.c_float), ("far_plane", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("frame_seed", C.c_uint32),
                ("view_id", C.c_uint32), ("max_variance", C.c_float), ("flags", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("gbuffer_rays", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("instance_visits", C.c_uint64), ("shadow_node_visits", C.c_uint64),
                ("shadow_tri_tests", C.c_uint64), ("shadow_instance_visits", C.c_uint64),
                ("samples", C.c_uint64), ("accumulated_rays", C.c_uint64),
                ("render_ms", C.c_double), ("cancelled", C.c_int32), ("denoise_fallback", C.c_int32),
                ("done_ms", C.c_double), ("spill_pushes", C.c_uint64), ("spill_pops", C.c_uint64),
                ("finish_rays", C.c_uint64), ("finish_node_visits", C.c_uint64), ("finish_tri_tests", C.c_uint64),
                ("finish_instance_visits", C.c_uint64)]

    @property
    def rays(self):
        return self.camera_rays + self.extension_rays + self.shadow_rays + self.gbuffer_rays


class Tuning(C.Structure):
    _fields_ = [("queue_paths", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("flags", C.c_uint32)]

    def as_tuple(self):
        return (self.queue_paths, self.tile_w, self.tile_h, self.flags)


class BuildStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("bvh_ms", C.c_float), ("binary_nodes", C.c_uint32),
                ("blas_nodes4", C.c_uint32), ("blas_bound", C.c_uint32), ("stack_bound", C.c_uint32),
                ("levels", C.c_uint32)]


class RebuildStats(C.Structure):
    _fields_ = [("device_ms", C.c_float), ("binary_depth", C.c_uint32), ("stack_bound", C.c_uint32),
                ("wide_nodes", C.c_uint32)]


class View(C.Structure):
    _fields_ = [("camera", Camera), ("params", RenderParams), ("d_pixels", C.c_void_p),
                ("d_sample_count", C.c_void_p), ("d_variance", C.c_void_p)]


class ViewHost(C.Structure):
    _fields_ = [("camera", Camera), ("params", RenderParams), ("pixels", C.c_void_p), ("sample_count", C.c_void_p),
                ("variance", C.c_void_p)]


# int (*)(void* user, const float* pixels, const uint32_t* sample_count, uint32_t w, uint32_t h, uint32_t samples_done)
PassCallback = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32,
                           C.c_uint32)


def params(width, height, spp, max_depth, tile=16, frame_seed=0, view_id=0, max_variance=0.0, flags=0):
    return RenderParams(width, height, spp, max_depth, tile, tile, frame_seed, view_id, max_variance, flags)


_lib = None


def _declare(lib):
    vp, u32, i32, f32, f64, u64, sz = C.c_void_p, C.c_uint32, C.c_int32, C.c_float, C.c_double, C.c_uint64, C.c_size_t
    P = C.POINTER
    sig = {
        "pico_hip_create": (i32, [i32, P(vp)]),
        "pico_hip_destroy": (None, [vp]),
        "pico_hip_last_error": (C.c_char_p, [vp]),
        "pico_hip_synchronize": (i32, [vp]),
        "pico_hip_device_name": (i32, [vp, C.c_char_p, sz]),
        "pico_hip_malloc": (i32, [vp, sz, P(vp)]),
        "pico_hip_free": (i32, [vp, vp]),
        "pico_hip_memcpy_h2d": (i32, [vp, vp, vp, sz]),
        "pico_hip_memcpy_d2h": (i32, [vp, vp, vp, sz]),
        "pico_hip_memset_d": (i32, [vp, vp, i32, sz]),
        "pico_hip_enable_kernel_timing": (i32, [vp, i32]),
        "pico_hip_kernel_time": (i32, [vp, C.c_char_p, P(u64), P(f64)]),
        "pico_hip_copy_d": (i32, [vp, vp, vp, sz]),
        "pico_hip_upload_scene": (i32, [vp, P(SceneDesc)]),
        "pico_hip_upload_scene_ex": (i32, [vp, P(SceneDesc), u32, P(BuildStats)]),
        "pico_hip_scene_info": (i32, [vp, P(u64), P(u64), P(u64), P(u32), P(u32)]),
        "pico_hip_set_instance_transforms": (i32, [vp, vp, P(RebuildStats)]),
        "pico_hip_set_instance_transforms_d": (i32, [vp, vp, P(RebuildStats)]),
        "pico_hip_closest_hit": (i32, [vp, u32, vp, vp, vp, vp, vp, vp]),
        "pico_hip_primary_ids": (i32, [vp, P(Camera), u32, u32, f32, f32, vp, vp]),
        "pico_hip_render": (i32, [vp, P(Camera), P(RenderParams), vp, vp, vp, vp, P(Stats)]),
        "pico_hip_render_d": (i32, [vp, P(Camera), P(RenderParams), vp, vp, vp, vp, P(Stats)]),
        "pico_hip_render_samples": (i32, [vp, P(Camera), P(RenderParams), u32, vp]),
        "pico_hip_render_accum64_d": (i32, [vp, P(Camera), P(RenderParams), u32, vp, P(Stats)]),
        "pico_hip_gbuffer": (i32, [vp, P(Camera), u32, u32, vp, vp, vp]),
        "pico_hip_gbuffer_d": (i32, [vp, P(Camera), u32, u32, vp, vp, vp]),
        "pico_hip_atrous": (i32, [vp, vp, vp, vp, vp, u32, u32, u32, u32, u32, vp]),
        "pico_hip_atrous_d": (i32, [vp, vp, vp, vp, vp, u32, u32, u32, u32, u32, vp, vp]),
        "pico_hip_tonemap": (i32, [vp, vp, u32, u32]),
        "pico_hip_atrous_fixed_d": (i32, [vp, vp, vp, vp, vp, u32, u32, u32, f32, f32, f32, f32, vp, vp]),
        "pico_hip_relmse_d": (i32, [vp, vp, vp, f64, u32, P(f64)]),
        "pico_hip_relmse_noalloc_d": (i32, [vp, vp, vp, f64, u32, vp]),
        "pico_hip_render_views_d": (i32, [vp, P(View), u32, u32, P(Stats)]),
        "pico_hip_render_views": (i32, [vp, P(ViewHost), u32, u32, P(Stats)]),
        "pico_hip_render_progressive": (i32, [vp, P(Camera), P(RenderParams), vp, vp, vp, PassCallback, vp, P(Stats)]),
        "pico_hip_set_tuning": (i32, [vp, P(Tuning)]),
        "pico_hip_get_tuning": (i32, [vp, P(Tuning)]),
        "pico_hip_render_views_denoise_d": (i32, [vp, P(View), u32, vp, P(vp), P(Stats)]),
        "pico_hip_autotune_views": (i32, [vp, P(View), u32, u32, P(Tuning), P(f32), P(f32)]),
        "pico_hip_cnn_create": (i32, [vp, u32, u32, u32, P(vp)]),
        "pico_hip_cnn_destroy": (None, [vp]),
        "pico_hip_cnn_num_params": (u32, []),
        "pico_hip_cnn_init": (i32, [vp, u32]),
        "pico_hip_cnn_set_params": (i32, [vp, vp]),
        "pico_hip_cnn_get_params": (i32, [vp, vp]),
        "pico_hip_cnn_denoise": (i32, [vp, vp, u32, u32, vp]),
        "pico_hip_cnn_denoise_d": (i32, [vp, vp, u32, u32, vp]),
        "pico_hip_cnn_inputs_d": (i32, [vp, vp, vp, vp, vp, u32, u32, vp]),
        "pico_hip_cnn_denoise_gbuffer": (i32, [vp, vp, vp, vp, vp, u32, u32, vp]),
        "pico_hip_cnn_denoise_gbuffer_d": (i32, [vp, vp, vp, vp, vp, u32, u32, vp]),
        "pico_hip_cnn_set_clip": (i32, [vp, f32]),
        "pico_hip_denoise_gbuffer_d": (i32, [vp, vp, vp, vp, vp, vp, u32, u32, u32, u32, vp, vp, vp]),
        "pico_hip_denoise_gbuffer": (i32, [vp, vp, vp, vp, vp, vp, u32, u32, u32, u32, vp, P(i32)]),
        "pico_hip_cnn_last_grad_norm": (i32, [vp, P(f64)]),
        "pico_hip_cnn_loss_grad": (i32, [vp, vp, vp, u32, u32, u32, P(f64), vp]),
        "pico_hip_cnn_train_step_d": (i32, [vp, vp, vp, u32, u32, u32, f32, P(f64)]),
        "pico_write_hdr": (i32, [C.c_char_p, u32, u32, vp, i32]),
        "pico_scene_cornell": (i32, [u32, u32, P(vp)]),
        "pico_scene_instanced": (i32, [u32, u32, u32, u32, u32, f32, u32, P(vp)]),
        "pico_scene_training": (i32, [u32, u32, P(vp)]),
        "pico_scene_cornell_glass": (i32, [u32, u32, P(vp)]),
        "pico_scene_shapes": (i32, [u32, u32, P(vp)]),
        "pico_scene_desc_get": (i32, [vp, P(SceneDesc)]),
        "pico_scene_load_json": (i32, [C.c_char_p, P(vp), C.c_char_p, sz]),
        "pico_scene_load_obj": (i32, [C.c_char_p, P(vp), C.c_char_p, sz]),
        "pico_scene_camera_count": (i32, [vp, P(u32)]),
        "pico_scene_camera_get": (i32, [vp, u32, C.c_char_p, sz, P(Camera)]),
        "pico_scene_default_camera": (i32, [vp, f32, P(Camera)]),
        "pico_scene_destroy": (None, [vp]),
        "pico_scene_set_sun": (i32, [vp, vp, vp]),
        "pico_scene_set_sky_image": (i32, [vp, P(SkyImage)]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def exported_symbols():
    """Every function name the C header declares (for the load check)."""
    hdr = os.path.join(os.path.dirname(_HERE), "include", "pico_hip.h")
    import re
    names = []
    with open(hdr) as f:
        for m in re.finditer(r"^\s*(?:int|void|uint32_t|const char\*)\s+\*?\s*(pico_\w+)\s*\(", f.read(), re.M):
            names.append(m.group(1))
    return names


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PicoError(PICO_ERR_STATE, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


class Scene:
    """A procedural scene (flattened pico_scene_desc owned by the library)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.desc = SceneDesc()
        rc = lib().pico_scene_desc_get(self._h, C.byref(self.desc))
        if rc:
            raise PicoError(rc, "pico_scene_desc_get")

    @staticmethod
    def cornell(seed=1, microfacet=MICROFACET_GGX):
        h = C.c_void_p()
        rc = lib().pico_scene_cornell(seed, microfacet, C.byref(h))
        if rc:
            raise PicoError(rc, "pico_scene_cornell")
        return Scene(h.value)

    @staticmethod
    def shapes(seed=1, microfacet=MICROFACET_GGX):
        h = C.c_void_p()
        rc = lib().pico_scene_shapes(seed, microfacet, C.byref(h))
        if rc:
            raise PicoError(rc, "pico_scene_shapes")
        return Scene(h.value)

    @staticmethod
    def cornell_glass(seed=1, microfacet=MICROFACET_GGX):
        h = C.c_void_p()
        rc = lib().pico_scene_cornell_glass(seed, microfacet, C.byref(h))
        if rc:
            raise PicoError(rc, "pico_scene_cornell_glass")
        return Scene(h.value)

    @staticmethod
    def instanced(seed=2, n_types=16, n_instances=1024, min_tris=2000, max_tris=20000, extent=100.0,
                  microfacet=MICROFACET_GGX):
        h = C.c_void_p()
        rc = lib().pico_scene_instanced(seed, n_types, n_instances, min_tris, max_tris, extent, microfacet, C.byref(h))
        if rc:
            raise PicoError(rc, "pico_scene_instanced")
        return Scene(h.value)

    @staticmethod
    def training(seed, microfacet=MICROFACET_GGX):
        h = C.c_void_p()
        rc = lib().pico_scene_training(seed, microfacet, C.byref(h))
        if rc:
            raise PicoError(rc, "pico_scene_training")
        return Scene(h.value)

    @staticmethod
    def load_json(path):
        """A scene file in the reference's JSON format (meshes from .obj files)."""
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().pico_scene_load_json(os.fsencode(path), C.byref(h), err, 1024)
        if rc:
            raise PicoError(rc, err.value.decode(errors="replace"))
        return Scene(h.value)

    @staticmethod
    def load_obj(path):
        """A model file (.obj) through the reference's import path: every mesh
        an instance, Light or Dielectric by its material's emission."""
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().pico_scene_load_obj(os.fsencode(path), C.byref(h), err, 1024)
        if rc:
            raise PicoError(rc, err.value.decode(errors="replace"))
        return Scene(h.value)

    def _refresh(self):
        rc = lib().pico_scene_desc_get(self._h, C.byref(self.desc))
        if rc:
            raise PicoError(rc, "pico_scene_desc_get")

    def set_sun(self, direction=None, colour=None):
        """The sun light (direction the light travels, colour), or None: off."""
        if direction is None:
            rc = lib().pico_scene_set_sun(self._h, None, None)
        else:
            d = (C.c_float * 3)(*direction)
            c = (C.c_float * 3)(*colour)
            rc = lib().pico_scene_set_sun(self._h, C.cast(d, C.c_void_p), C.cast(c, C.c_void_p))
        if rc:
            raise PicoError(rc, "pico_scene_set_sun")
        self._refresh()
        return self

    def set_sky_image(self, texels=None):
        """The sky image: uint8 [faces, h, w, 4] (RGBA8) or float32 (RGBA32F)
        with faces 6 (cube: +x, -x, +y, -y, +z, -z) or 1 (equirectangular); a
        3-D array is one face.  None restores the constant sky."""
        if texels is None:
            rc = lib().pico_scene_set_sky_image(self._h, None)
        else:
            t = np.ascontiguousarray(texels)
            if t.ndim == 3:
                t = t[None]
            if t.ndim != 4 or t.shape[-1] != 4 or t.dtype not in (np.uint8, np.float32):
                raise ValueError("sky texels must be [faces, h, w, 4] uint8 or float32")
            img = SkyImage(t.shape[2], t.shape[1], t.shape[0], TEXEL_RGBA32F if t.dtype == np.float32 else TEXEL_RGBA8,
                           t.ctypes.data)
            rc = lib().pico_scene_set_sky_image(self._h, C.byref(img))
        if rc:
            raise PicoError(rc, "pico_scene_set_sky_image")
        self._refresh()
        return self

    def cameras(self):
        """The named cameras of a loaded scene, {name: Camera}, in name order."""
        n = C.c_uint32()
        lib().pico_scene_camera_count(self._h, C.byref(n))
        out = {}
        for i in range(n.value):
            name = C.create_string_buffer(512)
            cam = Camera()
            lib().pico_scene_camera_get(self._h, i, name, 512, C.byref(cam))
            out[name.value.decode()] = cam
        return out

    def camera(self, aspect=1.0):
        cam = Camera()
        rc = lib().pico_scene_default_camera(self._h, aspect, C.byref(cam))
        if rc:
            raise PicoError(rc, "pico_scene_default_camera")
        return cam

    @property
    def n_triangles(self):
        d = self.desc
        offs = np.ctypeslib.as_array(C.cast(d.mesh_triangle_offset, C.POINTER(C.c_uint32)), (d.n_meshes + 1,))
        mesh = np.ctypeslib.as_array(C.cast(d.instance_mesh, C.POINTER(C.c_uint32)), (d.n_instances,))
        per = np.diff(offs.astype(np.int64))
        return int(per.sum()), int(per[mesh].sum())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.pico_scene_destroy(self._h)
            self._h = None


class DeviceArray:
    """Device memory owned by a Context (for the `_d` entry points)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx._check(lib().pico_hip_malloc(ctx._h, self.nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx._check(lib().pico_hip_memcpy_h2d(self.ctx._h, self.ptr, _ptr(a), a.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._check(lib().pico_hip_memcpy_d2h(self.ctx._h, _ptr(out), self.ptr, out.nbytes))
        return out

    def zero(self):
        self.ctx._check(lib().pico_hip_memset_d(self.ctx._h, self.ptr, 0, self.nbytes))

    def view(self, offset, nbytes):
        """A window [offset, offset + nbytes) of this allocation (not owned)."""
        assert 0 <= offset and offset + nbytes <= self.nbytes
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.nbytes, v._owner = self.ctx, int(nbytes), self
        v.ptr = C.c_void_p(self.ptr.value + int(offset))
        return v

    def free(self):
        if getattr(self, "_owner", None) is not None:
            self.ptr = None
            return
        if self.ptr is not None and self.ptr.value:
            lib().pico_hip_free(self.ctx._h, self.ptr)
        self.ptr = None


class Context:
    """One device context (stream, scene, work buffers)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().pico_hip_create(device, C.byref(h))
        if rc:
            raise PicoError(rc, "pico_hip_create failed (no GPU visible?)")
        self._h = h

    def _check(self, rc):
        if rc:
            raise PicoError(rc, lib().pico_hip_last_error(self._h).decode())
        return rc

    def close(self):
        if self._h:
            lib().pico_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_name(self):
        buf = C.create_string_buffer(256)
        self._check(lib().pico_hip_device_name(self._h, buf, 256))
        return buf.value.decode()

    def sync(self):
        self._check(lib().pico_hip_synchronize(self._h))

    def alloc(self, nbytes):
        return DeviceArray(self, nbytes)

    def upload_scene(self, scene, device_build=False):
        """Builds the BVHs (on the host, or with device_build on the GPU) and
        uploads the scene; returns BuildStats."""
        st = BuildStats()
        self._check(lib().pico_hip_upload_scene_ex(self._h, C.byref(scene.desc), BUILD_DEVICE if device_build else 0,
                                                   C.byref(st)))
        return st

    def scene_info(self):
        bn, bt, bi = C.c_uint64(), C.c_uint64(), C.c_uint64()
        td, bd = C.c_uint32(), C.c_uint32()
        self._check(lib().pico_hip_scene_info(self._h, C.byref(bn), C.byref(bt), C.byref(bi), C.byref(td), C.byref(bd)))
        return dict(bytes_nodes=bn.value, bytes_tris=bt.value, bytes_instances=bi.value, tlas_depth=td.value,
                    blas_depth=bd.value)

    def set_instance_transforms(self, transforms):
        """New column-major 4x4 transforms (n_instances x 16 f32, the scene
        description's layout) for every instance; the TLAS is rebuilt on the
        device.  Returns RebuildStats."""
        m = _f32(transforms, (-1, 16))
        st = RebuildStats()
        self._check(lib().pico_hip_set_instance_transforms(self._h, _ptr(m), C.byref(st)))
        return st

    def set_instance_transforms_d(self, d_transforms):
        st = RebuildStats()
        self._check(lib().pico_hip_set_instance_transforms_d(self._h, d_transforms.ptr, C.byref(st)))
        return st

    def closest_hit(self, origins, dirs):
        o, d = _f32(origins, (-1, 3)), _f32(dirs, (-1, 3))
        n = o.shape[0]
        t = np.empty(n, np.float32)
        inst = np.empty(n, np.uint32)
        prim = np.empty(n, np.uint32)
        uv = np.empty((n, 2), np.float32)
        self._check(lib().pico_hip_closest_hit(self._h, n, _ptr(o), _ptr(d), _ptr(t), _ptr(inst), _ptr(prim), _ptr(uv)))
        return t, inst, prim, uv

    def primary_ids(self, cam, w, h, xi=(0.0, 0.0)):
        inst = np.empty(w * h, np.uint32)
        prim = np.empty(w * h, np.uint32)
        self._check(lib().pico_hip_primary_ids(self._h, C.byref(cam), w, h, xi[0], xi[1], _ptr(inst), _ptr(prim)))
        return inst, prim

    def render(self, cam, p, pixels=None, counts=None, variance=None, cancel=None):
        n = p.width * p.height
        pixels = np.zeros((n, 3), np.float32) if pixels is None else pixels
        counts = np.zeros(n, np.uint32) if counts is None else counts
        variance = np.zeros((n, 3), np.float32) if variance is None else variance
        st = Stats()
        cptr = None if cancel is None else C.c_void_p(C.addressof(cancel))
        rc = lib().pico_hip_render(self._h, C.byref(cam), C.byref(p), _ptr(pixels), _ptr(counts), _ptr(variance), cptr,
                                   C.byref(st))
        if rc and rc != PICO_ERR_CANCELLED:
            self._check(rc)
        return pixels, counts, variance, st

    def render_progressive(self, cam, p, on_pass, pixels=None, counts=None, variance=None):
        """pico_hip_render with a snapshot after every chunk of samples:
        on_pass(pixels [H*W, 3], counts [H*W], samples_done) -> truthy to stop."""
        n = p.width * p.height
        pixels = np.zeros((n, 3), np.float32) if pixels is None else pixels
        counts = np.zeros(n, np.uint32) if counts is None else counts
        variance = np.zeros((n, 3), np.float32) if variance is None else variance

        def cb(_user, px, cnt, w, h, done):
            a = np.ctypeslib.as_array(px, (w * h * 3,)).reshape(w * h, 3).copy()
            b = np.ctypeslib.as_array(cnt, (w * h,)).copy()
            return 1 if on_pass(a, b, done) else 0
        fn = PassCallback(cb)
        st = Stats()
        rc = lib().pico_hip_render_progressive(self._h, C.byref(cam), C.byref(p), _ptr(pixels), _ptr(counts),
                                               _ptr(variance), fn, None, C.byref(st))
        if rc and rc != PICO_ERR_CANCELLED:
            self._check(rc)
        return pixels, counts, variance, st

    def render_views(self, cams, ps):
        """Host-buffer batched views: [(pixels, counts, variance)] per view."""
        arr = (ViewHost * len(ps))()
        outs = []
        for i, (cam, p) in enumerate(zip(cams, ps)):
            n = p.width * p.height
            o = (np.zeros((n, 3), np.float32), np.zeros(n, np.uint32), np.zeros((n, 3), np.float32))
            arr[i] = ViewHost(cam, p, o[0].ctypes.data, o[1].ctypes.data, o[2].ctypes.data)
            outs.append(o)
        st = (Stats * len(ps))()
        self._check(lib().pico_hip_render_views(self._h, arr, len(ps), 0, st))
        return outs, list(st)

    def render_d(self, cam, p, d_pixels, d_counts, d_variance):
        st = Stats()
        self._check(lib().pico_hip_render_d(self._h, C.byref(cam), C.byref(p), d_pixels.ptr, d_counts.ptr,
                                            d_variance.ptr, None, C.byref(st)))
        return st

    def render_samples(self, cam, p, sample_base=0):
        out = np.empty((p.spp, p.height * p.width, 3), np.float32)
        self._check(lib().pico_hip_render_samples(self._h, C.byref(cam), C.byref(p), sample_base, _ptr(out)))
        return out

    def render_accum64_d(self, cam, p, sample_base, d_sum):
        st = Stats()
        self._check(lib().pico_hip_render_accum64_d(self._h, C.byref(cam), C.byref(p), sample_base, d_sum.ptr,
                                                    C.byref(st)))
        return st

    def gbuffer(self, cam, w, h):
        n = np.empty((w * h, 3), np.float32)
        p = np.empty((w * h, 3), np.float32)
        a = np.empty((w * h, 3), np.float32)
        self._check(lib().pico_hip_gbuffer(self._h, C.byref(cam), w, h, _ptr(n), _ptr(p), _ptr(a)))
        return n, p, a

    def gbuffer_d(self, cam, w, h, dn, dp, da):
        self._check(lib().pico_hip_gbuffer_d(self._h, C.byref(cam), w, h, dn.ptr, dp.ptr, da.ptr))

    def atrous(self, rgb, normal, position, albedo, w, h, tile=(16, 16), levels=5):
        args = [_f32(x, (-1, 3)) for x in (rgb, normal, position, albedo)]
        out = np.empty((w * h, 3), np.float32)
        self._check(lib().pico_hip_atrous(self._h, *[_ptr(a) for a in args], w, h, tile[0], tile[1], levels, _ptr(out)))
        return out

    def atrous_d(self, d_rgb, d_n, d_p, d_a, w, h, d_scratch, d_out, tile=(16, 16), levels=5):
        self._check(lib().pico_hip_atrous_d(self._h, d_rgb.ptr, d_n.ptr, d_p.ptr, d_a.ptr, w, h, tile[0], tile[1], levels,
                                            d_scratch.ptr, d_out.ptr))

    def atrous_fixed_d(self, d_rgb, d_n, d_p, d_a, w, h, d_scratch, d_out, levels=5, sigmas=ATROUS_FIXED_SIGMAS):
        self._check(lib().pico_hip_atrous_fixed_d(self._h, d_rgb.ptr, d_n.ptr, d_p.ptr, d_a.ptr, w, h, levels, *sigmas,
                                                  d_scratch.ptr, d_out.ptr))

    def tonemap(self, rgb, w, h):
        out = _f32(rgb, (-1, 3)).copy()
        self._check(lib().pico_hip_tonemap(self._h, _ptr(out), w, h))
        return out

    def cnn_inputs_d(self, d_rgb, d_n, d_p, d_a, w, h, d_inputs):
        self._check(lib().pico_hip_cnn_inputs_d(self._h, d_rgb.ptr, d_n.ptr, d_p.ptr, d_a.ptr, w, h, d_inputs.ptr))

    def relmse_d(self, d_img, d_ref_sum, ref_count, n_pixels):
        out = C.c_double()
        self._check(lib().pico_hip_relmse_d(self._h, d_img.ptr, d_ref_sum.ptr, float(ref_count), n_pixels, C.byref(out)))
        return out.value

    def relmse_noalloc_d(self, d_img, d_ref_sum, ref_count, n_pixels, d_out):
        """The allocation-free relMSE: the mean goes to d_out (one device
        double) without a synchronisation."""
        self._check(lib().pico_hip_relmse_noalloc_d(self._h, d_img.ptr, d_ref_sum.ptr, float(ref_count), n_pixels,
                                                    d_out.ptr))

    def render_views_d(self, views, queue_paths=0):
        arr = (View * len(views))(*views)
        st = (Stats * len(views))()
        self._check(lib().pico_hip_render_views_d(self._h, arr, len(views), queue_paths, st))
        return list(st)

    def set_tuning(self, queue_paths=128 << 20, tile=(0, 0), flags=0):
        t = Tuning(queue_paths, tile[0], tile[1], flags)
        self._check(lib().pico_hip_set_tuning(self._h, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        self._check(lib().pico_hip_get_tuning(self._h, C.byref(t)))
        return t.as_tuple()

    def render_views_denoise_d(self, views, cnn, d_outs):
        arr = (View * len(views))(*views)
        outs = (C.c_void_p * len(views))(*[d.ptr.value for d in d_outs])
        st = (Stats * len(views))()
        self._check(lib().pico_hip_render_views_denoise_d(self._h, arr, len(views), cnn._h, outs, st))
        return list(st)

    def autotune_views(self, views, calib_spp=4):
        """Chosen (queue_paths, tile_w, tile_h) and the calibration ms of the
        choice and of the default launch shape."""
        arr = (View * len(views))(*views)
        t, cm, dm = Tuning(), C.c_float(), C.c_float()
        self._check(lib().pico_hip_autotune_views(self._h, arr, len(views), calib_spp, C.byref(t), C.byref(cm),
                                                  C.byref(dm)))
        return t.as_tuple(), cm.value, dm.value

    def denoise_gbuffer_d(self, cnn, d_rgb, d_n, d_p, d_a, w, h, d_scratch, d_out, d_flag, tile=(16, 16)):
        """Learned filter with the reference a-trous as fallback (cnn may be
        None); d_flag (one device u32) receives 1 when the fallback ran."""
        self._check(lib().pico_hip_denoise_gbuffer_d(self._h, None if cnn is None else cnn._h, d_rgb.ptr, d_n.ptr,
                                                     d_p.ptr, d_a.ptr, w, h, tile[0], tile[1], d_scratch.ptr, d_out.ptr,
                                                     d_flag.ptr))

    def denoise_gbuffer(self, cnn, rgb, normal, position, albedo, w, h, tile=(16, 16)):
        """Host variant: (output, used_fallback)."""
        args = [_f32(x, (-1, 3)) for x in (rgb, normal, position, albedo)]
        out = np.empty((w * h, 3), np.float32)
        fb = C.c_int32()
        self._check(lib().pico_hip_denoise_gbuffer(self._h, None if cnn is None else cnn._h, *[_ptr(a) for a in args], w,
                                                   h, tile[0], tile[1], _ptr(out), C.byref(fb)))
        return out, bool(fb.value)

    def enable_kernel_timing(self, on=True):
        self._check(lib().pico_hip_enable_kernel_timing(self._h, 1 if on else 0))

    def copy_d(self, dst, src, nbytes):
        self._check(lib().pico_hip_copy_d(self._h, dst.ptr, src.ptr, nbytes))

    def kernel_time(self, name):
        n, ms = C.c_uint64(), C.c_double()
        self._check(lib().pico_hip_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value


class CNN:
    """The learned kernel-predicting denoiser on the device (BF16 MFMA)."""

    def __init__(self, ctx, max_width, max_height, max_batch=1):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(lib().pico_hip_cnn_create(ctx._h, max_width, max_height, max_batch, C.byref(h)))
        self._h = h
        self.n_params = lib().pico_hip_cnn_num_params()

    def close(self):
        # a closed context has already destroyed the instances created on it
        if self._h and self.ctx._h:
            lib().pico_hip_cnn_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, seed):
        self.ctx._check(lib().pico_hip_cnn_init(self._h, seed))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self.ctx._check(lib().pico_hip_cnn_set_params(self._h, _ptr(p)))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self.ctx._check(lib().pico_hip_cnn_get_params(self._h, _ptr(p)))
        return p

    def denoise(self, inputs10, w, h):
        x = _f32(inputs10).reshape(10, h * w)
        out = np.empty((w * h, 3), np.float32)
        self.ctx._check(lib().pico_hip_cnn_denoise(self._h, _ptr(x), w, h, _ptr(out)))
        return out

    def denoise_d(self, d_in, w, h, d_out):
        self.ctx._check(lib().pico_hip_cnn_denoise_d(self._h, d_in.ptr, w, h, d_out.ptr))

    def denoise_gbuffer(self, rgb, normal, position, albedo, w, h):
        args = [_f32(x, (-1, 3)) for x in (rgb, normal, position, albedo)]
        out = np.empty((w * h, 3), np.float32)
        self.ctx._check(lib().pico_hip_cnn_denoise_gbuffer(self._h, *[_ptr(a) for a in args], w, h, _ptr(out)))
        return out

    def denoise_gbuffer_d(self, d_rgb, d_n, d_p, d_a, w, h, d_out):
        self.ctx._check(lib().pico_hip_cnn_denoise_gbuffer_d(self._h, d_rgb.ptr, d_n.ptr, d_p.ptr, d_a.ptr, w, h,
                                                            d_out.ptr))

    def load(self, path=None):
        """Parameters from a .npy file (the trained weights by default)."""
        p = np.load(path or weights_path(), allow_pickle=False)
        self.set_params(p)

    def save(self, path):
        np.save(path, self.get_params(), allow_pickle=False)

    def loss_grad(self, inputs, targets, batch, w, h):
        x = _f32(inputs).reshape(batch, 10, h * w)
        t = _f32(targets).reshape(batch, h * w, 3)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        self.ctx._check(lib().pico_hip_cnn_loss_grad(self._h, _ptr(x), _ptr(t), batch, w, h, C.byref(loss), _ptr(g)))
        return loss.value, g

    def set_clip(self, max_grad_norm):
        self.ctx._check(lib().pico_hip_cnn_set_clip(self._h, max_grad_norm))

    def last_grad_norm(self):
        out = C.c_double()
        self.ctx._check(lib().pico_hip_cnn_last_grad_norm(self._h, C.byref(out)))
        return out.value

    def train_step_d(self, d_in, d_tgt, batch, w, h, lr, want_loss=True):
        loss = C.c_double()
        self.ctx._check(lib().pico_hip_cnn_train_step_d(self._h, d_in.ptr, d_tgt.ptr, batch, w, h, lr,
                                                       C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None


# the network's convolutions: (input channels, output channels, taps, U-Net
# level); level l runs at 1/4^l of the pixels
CNN_LAYERS = [(10, 32, 9, 0), (32, 32, 9, 0), (32, 64, 9, 1), (64, 64, 9, 1), (64, 128, 9, 2), (128, 128, 9, 2),
              (192, 64, 9, 1), (64, 64, 9, 1), (96, 32, 9, 0), (32, 32, 9, 0), (32, 25, 1, 0)]


def cnn_flops_per_pixel():
    """Multiply-add FLOPs of one forward pass per full-resolution pixel."""
    return sum(2 * taps * cin * cout / 4 ** lvl for cin, cout, taps, lvl in CNN_LAYERS)


def write_hdr(path, rgb, width, height, reverse=True):
    """Radiance .hdr file of a packed RGB image; `reverse` reverses the pixel
    order end to end first, as Scene::render_scene_to_file does."""
    a = _f32(rgb, (-1, 3))
    rc = lib().pico_write_hdr(os.fsencode(path), width, height, _ptr(a), 1 if reverse else 0)
    if rc:
        raise PicoError(rc, f"pico_write_hdr({path})")


def weights_path():
    """The trained denoiser's parameters shipped with the package."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "kpn.npy")
