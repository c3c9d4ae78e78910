"""Python binding of the MI355X BDPT render core (libomv.so, C-ABI in include/omv.h).

The classes mirror the reference pathtracer's host interfaces one to one:

  Scene     RenderScene   (pathtracer/src/scene/RenderScene.h:109-204)
  Camera    PinholeCamera (pathtracer/src/sensors/PinholeCamera.h)
  Renderer  BDPTRenderer  (pathtracer/src/renderer/BDPTRenderer.h:147-155)

Only the compiled HIP library is used; importing this module fails loudly when
libomv.so is missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libomv.so"

OMV_OK = 0
EXT_NONE = 0
EXT_OM_ROUGH = 1
EXT_HDRI_NEE = 2
BUILD_HOST_SAH, BUILD_DEVICE_LBVH = 0, 1
EXT_LIGHT_SPLAT = 4

FC_SUBJECTS = ("pid", "gid", "mid", "rmp", "mmp", "md", "sg", "mn", "in")
PBR_BASE_COLOR, PBR_METALLIC_ROUGHNESS, PBR_NORMAL, PBR_EMISSIVITY, PBR_OCCLUSION = range(5)
ALPHA_BLEND, ALPHA_MASK, ALPHA_OPAQUE = range(3)
SCHEDULE_WAVEFRONT, SCHEDULE_PER_SAMPLE = 0, 1


class RendererParams(C.Structure):
    _fields_ = [
        ("camera_bounces", C.c_int), ("light_bounces", C.c_int), ("max_path_length", C.c_int),
        ("samples_per_pixel", C.c_int), ("light_distance_attenuation", C.c_int),
        ("light_scale_value", C.c_float), ("light_min_distance", C.c_float), ("ray_cut_pix_value", C.c_float),
        ("gamma", C.c_float), ("number_of_cores", C.c_int), ("false_color_mode", C.c_int),
        ("seed", C.c_uint32), ("extensions", C.c_uint32), ("om_roughness", C.c_float),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("position", C.c_float * 3), ("rotation", C.c_float * 9),
        ("half_w", C.c_float), ("half_h", C.c_float), ("focus", C.c_float),
    ]


class PbrDesc(C.Structure):
    _fields_ = [
        ("base_color_factor", C.c_float * 4), ("emissive_factor", C.c_float * 3), ("metallic_factor", C.c_float),
        ("roughness_factor", C.c_float), ("normal_scale", C.c_float), ("occlusion_strength", C.c_float),
        ("ior", C.c_float), ("alpha_cutoff", C.c_float), ("alpha_mode", C.c_int), ("double_sided", C.c_int),
        ("texture_map", C.c_int * 5), ("texture_channel", C.c_int * 5), ("texture_transform", (C.c_float * 9) * 5),
    ]


class IorTable(C.Structure):
    _fields_ = [
        ("temperature", C.c_float), ("n_count", C.c_int), ("n_wavelength", C.POINTER(C.c_float)),
        ("n_value", C.POINTER(C.c_float)), ("k_count", C.c_int), ("k_wavelength", C.POINTER(C.c_float)),
        ("k_value", C.POINTER(C.c_float)),
    ]


class RenderTuning(C.Structure):
    _fields_ = [("trace_keep", C.c_int), ("trace_blocks_per_cu", C.c_int), ("chunk_samples", C.c_int), ("traversal", C.c_int),
                ("sort_rays", C.c_int), ("chunk_streams", C.c_int), ("chunk_stagger", C.c_int), ("occlusion", C.c_int),
                ("xcd_bands", C.c_int), ("tail_round", C.c_int)]


class AdjointTargets(C.Structure):
    _fields_ = [("n_textures", C.c_int), ("texture_ids", C.POINTER(C.c_int)), ("texture_grads", C.POINTER(C.c_void_p)),
                ("n_om_materials", C.c_int), ("om_material_ids", C.POINTER(C.c_int)), ("om_grads", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int), ("tile_waves", C.c_int), ("use_graph", C.c_int), ("streams", C.c_int)]


class RenderStats(C.Structure):
    _fields_ = [
        ("samples", C.c_uint64), ("rays", C.c_uint64), ("camera_rays", C.c_uint64), ("light_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64), ("nodes_visited", C.c_uint64), ("tris_tested", C.c_uint64),
        ("stack_overflows", C.c_uint64), ("render_ms", C.c_double), ("kernel_ms", C.c_double),
        ("trace_ms", C.c_double), ("trace_launches", C.c_uint64),
    ]


def pbr_desc(base_color=(1.0, 1.0, 1.0, 1.0), metallic=1.0, roughness=1.0, emissive=(0.0, 0.0, 0.0), normal_scale=1.0,
             occlusion_strength=1.0, ior=1.5, alpha_cutoff=0.5, alpha_mode=ALPHA_OPAQUE, double_sided=False,
             textures=None, channels=None, transforms=None) -> PbrDesc:
    """PbrMaterial defaults (PbrMaterial.cpp:104-120); textures maps slot -> texture id."""
    d = PbrDesc()
    d.base_color_factor[:] = [float(x) for x in base_color]
    d.emissive_factor[:] = [float(x) for x in emissive]
    d.metallic_factor, d.roughness_factor = float(metallic), float(roughness)
    d.normal_scale, d.occlusion_strength, d.ior, d.alpha_cutoff = float(normal_scale), float(occlusion_strength), float(ior), float(alpha_cutoff)
    d.alpha_mode, d.double_sided = int(alpha_mode), int(bool(double_sided))
    for s in range(5):
        d.texture_map[s] = -1
        d.texture_channel[s] = 0
        d.texture_transform[s][:] = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    for s, t in (textures or {}).items():
        d.texture_map[s] = int(t)
    for s, ch in (channels or {}).items():
        d.texture_channel[s] = int(ch)
    for s, m in (transforms or {}).items():
        d.texture_transform[s][:] = [float(x) for x in np.asarray(m, np.float32).reshape(9)]
    return d


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def ior_tables(tables):
    """tables: list of dicts {temperature, n: [[wl, v]...], k: [[wl, v]...]} -> (array, keepalive)."""
    arr = (IorTable * len(tables))()
    keep = []
    for i, t in enumerate(tables):
        n = np.asarray(t["n"], np.float32).reshape(-1, 2)
        k = np.asarray(t["k"], np.float32).reshape(-1, 2)
        nw, nv = np.ascontiguousarray(n[:, 0]), np.ascontiguousarray(n[:, 1])
        kw, kv = np.ascontiguousarray(k[:, 0]), np.ascontiguousarray(k[:, 1])
        keep += [nw, nv, kw, kv]
        arr[i].temperature = float(t["temperature"])
        arr[i].n_count, arr[i].k_count = len(nw), len(kw)
        arr[i].n_wavelength, arr[i].n_value = _ptr(nw, C.c_float), _ptr(nv, C.c_float)
        arr[i].k_wavelength, arr[i].k_value = _ptr(kw, C.c_float), _ptr(kv, C.c_float)
    return arr, keep


class SceneApi:
    """Scene set-up shared by the product and the CPU oracle bindings (same
    entry points, different prefix)."""

    prefix = "omv"

    def _fn(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{self.prefix}_{what} failed ({rc}): {self._err()}")

    def _err(self):
        return ""

    def allocate(self, meshes, instances, materials, textures, lights):
        self._check(self._fn("scene_allocate")(self.h, meshes, instances, materials, textures, lights), "scene_allocate")

    def set_mesh(self, mesh_id, vertices, faces, mat_id, normals=None, tangents=None, uvs=(), tex_n=None, tangent_stride=4):
        """RenderScene::setMesh.  tangents are 4 floats per vertex read at
        tangent_stride floats apart (main.cpp:386 passes glTF VEC4 tangents
        with stride 3, which this reproduces when asked)."""
        v = _f32(vertices, (-1, 3))
        f = np.ascontiguousarray(np.asarray(faces, np.uint32).reshape(-1, 3))
        n = _f32(normals, (-1, 3)) if normals is not None else None
        t = None
        if tangents is not None:
            t = _f32(tangents).reshape(-1)
            if len(v) and t.size < tangent_stride * (len(v) - 1) + 4:
                raise ValueError(f"set_mesh: {t.size} tangent floats for {len(v)} vertices at stride {tangent_stride}")
        uvl = [(_f32(u, (-1, 2)) if u is not None else None) for u in uvs]
        texn = len(uvl) if tex_n is None else tex_n
        arr = (C.POINTER(C.c_float) * max(1, texn))()
        for i in range(texn):
            arr[i] = _ptr(uvl[i], C.c_float) if i < len(uvl) and uvl[i] is not None else None
        rc = self._fn("scene_set_mesh")(self.h, mesh_id, len(v), len(f), texn, mat_id, _ptr(f, C.c_uint32), _ptr(v, C.c_float), 3,
                                         _ptr(n, C.c_float), 3, _ptr(t, C.c_float), int(tangent_stride), arr, 2)
        self._check(rc, "scene_set_mesh")

    def set_instance(self, inst_id, trans, mesh_id):
        m = _f32(trans, (16,))
        self._check(self._fn("scene_set_instance")(self.h, inst_id, _ptr(m, C.c_float), mesh_id), "scene_set_instance")

    def set_material_om(self, mat_id, tables, temperature=300.0):
        arr, keep = ior_tables(tables)
        self._check(self._fn("scene_set_material_om")(self.h, mat_id, C.c_float(temperature), len(tables), arr), "scene_set_material_om")
        del keep

    def set_material_om_nk(self, mat_id, n, k):
        nn, kk = _f32(n, (3,)), _f32(k, (3,))
        self._check(self._fn("scene_set_material_om_nk")(self.h, mat_id, _ptr(nn, C.c_float), _ptr(kk, C.c_float)), "scene_set_material_om_nk")

    def set_material_pbr(self, mat_id, desc: PbrDesc):
        self._check(self._fn("scene_set_material_pbr")(self.h, mat_id, C.byref(desc)), "scene_set_material_pbr")

    def set_texture(self, tex_id, rgba):
        a = np.ascontiguousarray(np.asarray(rgba, np.uint8))
        h, w = a.shape[0], a.shape[1]
        self._check(self._fn("scene_set_texture")(self.h, tex_id, w, h, _ptr(a, C.c_uint8)), "scene_set_texture")

    def set_texture_f32(self, tex_id, rgba):
        a = _f32(rgba)
        h, w = a.shape[0], a.shape[1]
        self._check(self._fn("scene_set_texture_f32")(self.h, tex_id, w, h, _ptr(a, C.c_float)), "scene_set_texture_f32")

    def set_light_point(self, light_id, pos, intensity, range_=-1.0):
        p, i = _f32(pos, (3,)), _f32(intensity, (3,))
        self._check(self._fn("scene_set_light_point")(self.h, light_id, _ptr(p, C.c_float), _ptr(i, C.c_float), C.c_float(range_)),
                    "scene_set_light_point")

    def set_background_hdr(self, rgb, rotation, scale):
        a = _f32(rgb)
        h, w = a.shape[0], a.shape[1]
        r = _f32(rotation, (9,))
        self._check(self._fn("scene_set_background_hdr")(self.h, w, h, _ptr(a, C.c_float), _ptr(r, C.c_float), C.c_float(scale)),
                    "scene_set_background_hdr")

    def set_background_constant(self, rgb):
        a = _f32(rgb, (3,))
        self._check(self._fn("scene_set_background_constant")(self.h, _ptr(a, C.c_float)), "scene_set_background_constant")

    def set_build_mode(self, mode: int):
        """omv_scene_set_build_mode: BUILD_HOST_SAH (default) or BUILD_DEVICE_LBVH."""
        self._check(self._fn("scene_set_build_mode")(self.h, int(mode)), "scene_set_build_mode")

    def set_false_color_subject(self, name):
        self._check(self._fn("scene_set_false_color_subject")(self.h, name.encode()), "scene_set_false_color_subject")

    def commit(self):
        buf = C.create_string_buffer(1 << 16)
        rc = self._fn("scene_commit")(self.h, buf, len(buf))
        self._check(rc, "scene_commit")
        return buf.value.decode()

    def mesh_vertices(self, mesh_id):
        vn, fn, st, mat = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._check(self._fn("scene_mesh_info")(self.h, mesh_id, C.byref(vn), C.byref(fn), C.byref(st), C.byref(mat)), "scene_mesh_info")
        out = np.zeros(max(1, vn.value * st.value), np.float32)
        self._check(self._fn("scene_get_vertices")(self.h, mesh_id, _ptr(out, C.c_float), out.size), "scene_get_vertices")
        return out[: vn.value * st.value].reshape(vn.value, st.value), mat.value


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"libomv.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = C.CDLL(str(LIB_PATH))
    lib.omv_last_error.restype = C.c_char_p
    lib.omv_version.restype = C.c_char_p
    vp = C.c_void_p
    for name, args in {
        "omv_scene_create": [C.POINTER(vp)], "omv_scene_destroy": [vp],
        "omv_scene_allocate": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
        "omv_scene_set_mesh": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.c_int,
                               C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.POINTER(C.c_float)), C.c_int],
        "omv_scene_set_instance": [vp, C.c_int, C.POINTER(C.c_float), C.c_int],
        "omv_scene_set_material_om": [vp, C.c_int, C.c_float, C.c_int, C.POINTER(IorTable)],
        "omv_scene_set_material_om_nk": [vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "omv_scene_set_material_pbr": [vp, C.c_int, C.POINTER(PbrDesc)],
        "omv_scene_set_texture": [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8)],
        "omv_scene_set_texture_f32": [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)],
        "omv_scene_set_light_point": [vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float],
        "omv_scene_set_background_hdr": [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float],
        "omv_scene_set_background_constant": [vp, C.POINTER(C.c_float)],
        "omv_scene_set_false_color_subject": [vp, C.c_char_p], "omv_scene_set_build_mode": [vp, C.c_int],
        "omv_scene_commit": [vp, C.c_char_p, C.c_size_t],
        "omv_scene_mesh_info": [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "omv_scene_get_vertices": [vp, C.c_int, C.POINTER(C.c_float), C.c_size_t],
        "omv_scene_bvh_info": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_size_t)],
        "omv_scene_env_tables": [vp, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_float)],
        "omv_camera_create": [C.POINTER(vp), C.c_int, C.c_int], "omv_camera_destroy": [vp],
        "omv_camera_set_viewpoint": [vp] + [C.POINTER(C.c_float)] * 4,
        "omv_camera_adjust_to_resolution": [vp], "omv_camera_set_yfov": [vp, C.c_float],
        "omv_camera_set_real_height": [vp, C.c_float, C.c_float], "omv_camera_set_focus": [vp, C.c_float],
        "omv_save_image": [C.c_char_p, C.POINTER(C.c_float), C.c_int, C.c_int],
        "omv_fc_color_table": [C.c_uint32, C.c_int, C.c_uint32, C.POINTER(C.c_uint32)],
        "omv_encode_image": [C.c_char_p, C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_size_t,
                             C.POINTER(C.c_size_t)],
        "omv_camera_load_properties": [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "omv_camera_get_desc": [vp, C.POINTER(CameraDesc)],
        "omv_camera_get_impression": [vp, C.c_int, C.c_int, C.POINTER(C.c_float)],
        "omv_camera_get_film": [vp, C.POINTER(C.c_float), C.c_size_t],
        "omv_renderer_create": [C.POINTER(vp)], "omv_renderer_destroy": [vp],
        "omv_renderer_set_parameters": [vp, C.POINTER(RendererParams)], "omv_renderer_set_scene": [vp, vp],
        "omv_renderer_render": [vp, vp], "omv_renderer_render_device": [vp, vp, C.c_void_p], "omv_renderer_render_noalloc": [vp, vp, C.c_void_p, C.c_void_p, C.c_void_p], "omv_renderer_render_fc": [vp, vp, C.POINTER(C.c_int32), C.POINTER(C.c_uint8)],
        "omv_renderer_get_stats": [vp, C.POINTER(RenderStats)],
        "omv_renderer_backward": [vp, vp, C.c_void_p, C.POINTER(AdjointTargets)],
        "omv_renderer_render_detached": [vp, vp, vp, C.c_void_p],
        "omv_renderer_post": [vp, vp, C.c_float, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_uint8)],
        "omv_scene_update_texture": [vp, C.c_int, C.c_void_p],
        "omv_scene_update_material_om_nk": [vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "omv_renderer_set_schedule": [vp, C.c_int], "omv_renderer_get_schedule": [vp, C.POINTER(C.c_int)],
        "omv_renderer_set_tuning": [vp, C.POINTER(RenderTuning)], "omv_renderer_get_tuning": [vp, C.POINTER(RenderTuning)],
        "omv_renderer_params_default": [C.POINTER(RendererParams)],
        "omv_batch_create": [C.POINTER(vp)], "omv_batch_destroy": [vp],
        "omv_batch_add_job": [vp, vp, vp, C.POINTER(RendererParams), C.POINTER(C.c_int)],
        "omv_batch_job_count": [vp, C.POINTER(C.c_int)],
        "omv_batch_set_tuning": [vp, C.POINTER(BatchTuning)], "omv_batch_get_tuning": [vp, C.POINTER(BatchTuning)],
        "omv_batch_autotune": [vp, C.c_int, C.POINTER(BatchTuning)], "omv_batch_run": [vp],
        "omv_batch_get_film": [vp, C.c_int, C.POINTER(C.c_float), C.c_size_t],
        "omv_batch_get_stats": [vp, C.POINTER(RenderStats)],
        "omv_device_synchronize": [], "omv_device_info": [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)],
    }.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int if name != "omv_renderer_params_default" else None
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def device_synchronize():
    rc = lib().omv_device_synchronize()
    if rc != 0:
        raise RuntimeError(lib().omv_last_error().decode())


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cus, mem = C.c_int(), C.c_size_t()
    rc = lib().omv_device_info(name, len(name), C.byref(cus), C.byref(mem))
    if rc != 0:
        raise RuntimeError(lib().omv_last_error().decode())
    return {"arch": name.value.decode(), "compute_units": cus.value, "memory_bytes": mem.value}


def default_params(**kw) -> RendererParams:
    p = RendererParams()
    lib().omv_renderer_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Scene(SceneApi):
    prefix = "omv"

    def update_texture(self, tex_id: int, texels):
        """Replace a committed texture's texels in place: a numpy array (host)
        or an int device address (e.g. a torch tensor's data_ptr())."""
        if isinstance(texels, int):
            src = C.c_void_p(texels)
        else:
            arr = np.ascontiguousarray(texels)
            self._keep = arr
            src = C.c_void_p(arr.ctypes.data)
        rc = self.lib.omv_scene_update_texture(self.h, tex_id, src)
        if rc != 0:
            raise RuntimeError(self.lib.omv_last_error().decode())

    def update_om(self, mat_id: int, n, k):
        nn, kk = _f32(n, (3,)), _f32(k, (3,))
        rc = self.lib.omv_scene_update_material_om_nk(self.h, mat_id, _ptr(nn, C.c_float), _ptr(kk, C.c_float))
        if rc != 0:
            raise RuntimeError(self.lib.omv_last_error().decode())

    def __init__(self):
        self.lib = lib()
        h = C.c_void_p()
        self._check(self.lib.omv_scene_create(C.byref(h)), "scene_create")
        self.h = h

    def _err(self):
        return self.lib.omv_last_error().decode()

    def bvh_info(self):
        n, d, r, b = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        self._check(self.lib.omv_scene_bvh_info(self.h, C.byref(n), C.byref(d), C.byref(r), C.byref(b)), "scene_bvh_info")
        return {"nodes": n.value, "depth": d.value, "tri_records": r.value, "device_bytes": b.value}

    def env_tables(self, w: int, h: int):
        """The environment sampling tables built on the device at commit
        (OMV_EXT_HDRI_NEE): row prob/alias (h), column prob/alias and texel
        probabilities (h, w)."""
        rp, ra = np.zeros(h, np.float32), np.zeros(h, np.int32)
        cp, ca, tp = np.zeros((h, w), np.float32), np.zeros((h, w), np.int32), np.zeros((h, w), np.float32)
        self._check(self.lib.omv_scene_env_tables(self.h, _ptr(rp, C.c_float), _ptr(ra, C.c_int32), _ptr(cp, C.c_float),
                                                  _ptr(ca, C.c_int32), _ptr(tp, C.c_float)), "scene_env_tables")
        return rp, ra, cp, ca, tp

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.omv_scene_destroy(self.h)
            self.h = None


class Camera:
    """PinholeCamera: viewpoint from (position, forward, up) as main.cpp:459-476 builds it."""

    def __init__(self, width, height, position=(0, 0, 0), forward=(0, 0, -1), up=(0, 1, 0), yfov=60.0):
        self.lib = lib()
        h = C.c_void_p()
        self._check(self.lib.omv_camera_create(C.byref(h), width, height))
        self.h = h
        self.width, self.height = int(width), int(height)
        self.look(position, forward, up)
        self._check(self.lib.omv_camera_adjust_to_resolution(self.h))
        self._check(self.lib.omv_camera_set_yfov(self.h, C.c_float(yfov)))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.omv_last_error().decode())

    def look(self, position, forward, up):
        fwd = np.asarray(forward, np.float32)
        upv = np.asarray(up, np.float32)
        n = np.float32(np.sqrt(np.float32(fwd[0] * fwd[0] + fwd[1] * fwd[1]) + np.float32(fwd[2] * fwd[2])))
        fwd = (fwd * (np.float32(1.0) / n)).astype(np.float32)
        left = _cross32(fwd, upv)
        upv = _cross32(left, fwd)
        self.set_viewpoint(position, left, upv, fwd)

    def set_viewpoint(self, pos, x, y, z):
        a = [_f32(v, (3,)) for v in (pos, x, y, z)]
        self._check(self.lib.omv_camera_set_viewpoint(self.h, *[_ptr(v, C.c_float) for v in a]))

    def desc(self) -> CameraDesc:
        d = CameraDesc()
        self._check(self.lib.omv_camera_get_desc(self.h, C.byref(d)))
        return d

    def load_properties(self, aspect=None, focus=None, yfov=None):
        """PinholeCamera::loadProperties with a camera file's values
        (assets.read_camera_properties); None = key absent."""
        vals = [None if v is None else C.c_float(float(v)) for v in (aspect, focus, yfov)]
        self._check(self.lib.omv_camera_load_properties(self.h, *[C.byref(v) if v is not None else None for v in vals]))

    def film(self):
        out = np.empty(self.width * self.height * 4, np.float32)
        self._check(self.lib.omv_camera_get_film(self.h, _ptr(out, C.c_float), out.size))
        return out.reshape(self.height, self.width, 4)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.omv_camera_destroy(self.h)
            self.h = None


def _cross32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], np.float32)


class Renderer:
    def __init__(self, params: RendererParams | None = None):
        self.lib = lib()
        h = C.c_void_p()
        self._check(self.lib.omv_renderer_create(C.byref(h)))
        self.h = h
        self.params = params if params is not None else default_params()
        self.set_parameters(self.params)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.omv_last_error().decode())

    def set_parameters(self, p: RendererParams):
        self.params = p
        self._check(self.lib.omv_renderer_set_parameters(self.h, C.byref(p)))

    def set_scene(self, scene: Scene):
        self._scene = scene
        self._check(self.lib.omv_renderer_set_scene(self.h, scene.h))

    def set_schedule(self, schedule: int):
        """SCHEDULE_WAVEFRONT (default) or SCHEDULE_PER_SAMPLE; same film either way."""
        self._check(self.lib.omv_renderer_set_schedule(self.h, schedule))

    def schedule(self) -> int:
        v = C.c_int()
        self._check(self.lib.omv_renderer_get_schedule(self.h, C.byref(v)))
        return v.value

    def set_tuning(self, **kw):
        """trace_keep, trace_blocks_per_cu, chunk_samples, traversal, sort_rays, chunk_streams, chunk_stagger,
        occlusion, xcd_bands, tail_round (the film does not depend on them)."""
        t = RenderTuning()
        self._check(self.lib.omv_renderer_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(self.lib.omv_renderer_set_tuning(self.h, C.byref(t)))

    def tuning(self) -> dict:
        t = RenderTuning()
        self._check(self.lib.omv_renderer_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in RenderTuning._fields_}

    def render(self, cam: Camera):
        self._check(self.lib.omv_renderer_render(self.h, cam.h))
        return cam.film()

    def render_device(self, cam: Camera, film_ptr: int):
        """render() with the film written to a device buffer (film_ptr = device
        address of H*W*4 floats) and left there."""
        self._check(self.lib.omv_renderer_render_device(self.h, cam.h, C.c_void_p(film_ptr)))

    def render_noalloc(self, cam: Camera, film_ptr: int, stream: int = 0, stats: bool = True):
        """Allocation-free render into a caller-owned device film (film_ptr =
        device address of H*W*4 floats) on `stream` (0 = the null stream).
        stats=True waits and returns the counters; False only enqueues."""
        st = RenderStats() if stats else None
        self._check(self.lib.omv_renderer_render_noalloc(self.h, cam.h, C.c_void_p(film_ptr), C.c_void_p(stream or None),
                                                         C.byref(st) if st is not None else None))
        return {k: getattr(st, k) for k, _ in RenderStats._fields_} if st is not None else None

    def backward(self, cam: Camera, d_film_ptr: int, textures=(), om_materials=(), om_grads_ptr: int = 0):
        """Path-replay adjoint of the last render: d_film_ptr = device address of
        d loss / d film (H*W*4 floats); textures = [(tex_id, grad device address)];
        om_materials = [mat_id] with om_grads_ptr -> 6 floats each.  Adds."""
        t = AdjointTargets()
        ids = (C.c_int * max(1, len(textures)))(*[i for i, _ in textures])
        ptrs = (C.c_void_p * max(1, len(textures)))(*[g for _, g in textures])
        oms = (C.c_int * max(1, len(om_materials)))(*om_materials)
        t.n_textures, t.texture_ids, t.texture_grads = len(textures), ids, ptrs
        t.n_om_materials, t.om_material_ids, t.om_grads = len(om_materials), oms, C.c_void_p(om_grads_ptr or None)
        self._check(self.lib.omv_renderer_backward(self.h, cam.h, C.c_void_p(d_film_ptr), C.byref(t)))

    def render_detached(self, cam: Camera, eval_scene: "Scene") -> np.ndarray:
        """FP64 film of the functional backward() differentiates: paths
        sampled from this renderer's scene, weights with eval_scene's
        parameters.  Returns (H, W, 4) float64."""
        out = np.zeros(cam.width * cam.height * 4, np.float64)
        self._check(self.lib.omv_renderer_render_detached(self.h, cam.h, eval_scene.h, C.c_void_p(out.ctypes.data)))
        return out.reshape(cam.height, cam.width, 4)

    def post(self, cam: Camera, gamma=0.5, denoise=False):
        """main.cpp's post pass on the device: (gamma-corrected [+ median] float film, 8-bit image)."""
        n = cam.width * cam.height * 4
        f = np.zeros(n, np.float32)
        u = np.zeros(n, np.uint8)
        self._check(self.lib.omv_renderer_post(self.h, cam.h, C.c_float(gamma), int(bool(denoise)), _ptr(f, C.c_float), _ptr(u, C.c_uint8)))
        return f.reshape(cam.height, cam.width, 4), u.reshape(cam.height, cam.width, 4)

    def render_fc(self, cam: Camera):
        npix = cam.width * cam.height
        ids = np.zeros(4 * npix, np.int32)
        tie = np.zeros(npix, np.uint8)
        self._check(self.lib.omv_renderer_render_fc(self.h, cam.h, _ptr(ids, C.c_int32), _ptr(tie, C.c_uint8)))
        return cam.film(), ids.reshape(4, cam.height, cam.width), tie.reshape(cam.height, cam.width)

    def stats(self) -> dict:
        s = RenderStats()
        self._check(self.lib.omv_renderer_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in RenderStats._fields_}

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.omv_renderer_destroy(self.h)
            self.h = None


class Batch:
    """Many render jobs (scene, camera, parameters) in one persistent launch.

    Each job's film equals Renderer.render's for the same inputs.  Scenes must
    stay alive and committed while the batch exists; cameras are copied when
    the job is added."""

    def __init__(self):
        self.lib = lib()
        h = C.c_void_p()
        self._check(self.lib.omv_batch_create(C.byref(h)))
        self.h = h
        self._jobs = []      # (scene, width, height): keeps scenes alive

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.omv_last_error().decode())

    def add_job(self, scene: Scene, cam: Camera, params: RendererParams) -> int:
        jid = C.c_int()
        self._check(self.lib.omv_batch_add_job(self.h, scene.h, cam.h, C.byref(params), C.byref(jid)))
        self._jobs.append((scene, cam.width, cam.height))
        return jid.value

    def __len__(self):
        return len(self._jobs)

    def set_tuning(self, blocks_per_cu=0, tile_waves=1, use_graph=True, streams=0):
        t = BatchTuning(blocks_per_cu, tile_waves, 1 if use_graph else 0, streams)
        self._check(self.lib.omv_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self) -> dict:
        t = BatchTuning()
        self._check(self.lib.omv_batch_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in BatchTuning._fields_}

    def autotune(self, reps=1) -> dict:
        t = BatchTuning()
        self._check(self.lib.omv_batch_autotune(self.h, reps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in BatchTuning._fields_}

    def run(self):
        self._check(self.lib.omv_batch_run(self.h))

    def film(self, job: int) -> np.ndarray:
        _, w, h = self._jobs[job]
        out = np.empty(w * h * 4, np.float32)
        self._check(self.lib.omv_batch_get_film(self.h, job, _ptr(out, C.c_float), out.size))
        return out.reshape(h, w, 4)

    def stats(self) -> dict:
        s = RenderStats()
        self._check(self.lib.omv_batch_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in RenderStats._fields_}

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.omv_batch_destroy(self.h)
            self.h = None


def save_image(filename: str, rgba) -> None:
    """saveImage of the reference main (main.cpp:86-120): the (H, W, 4)
    gamma-corrected float image as .png/.bmp/.tga/.jpg (8-bit, 255*f
    truncated) or .pfm, rows written bottom-up, bytes as stb_image_write's."""
    a = _f32(rgba)
    h, w = a.shape[:2]
    rc = lib().omv_save_image(str(filename).encode(), _ptr(a, C.c_float), int(w), int(h))
    if rc != 0:
        raise RuntimeError(f"omv_save_image failed ({rc}): {lib().omv_last_error().decode()}")


def encode_image(ext: str, rgba) -> bytes:
    """The bytes save_image writes for file extension ``ext``."""
    a = _f32(rgba)
    h, w = a.shape[:2]
    n = C.c_size_t()
    rc = lib().omv_encode_image(ext.encode(), _ptr(a, C.c_float), int(w), int(h), None, 0, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"omv_encode_image failed ({rc}): {lib().omv_last_error().decode()}")
    out = (C.c_uint8 * n.value)()
    rc = lib().omv_encode_image(ext.encode(), _ptr(a, C.c_float), int(w), int(h), out, n.value, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"omv_encode_image failed ({rc}): {lib().omv_last_error().decode()}")
    return bytes(out)


def fc_color_table(seed: int, subject: int, size: int) -> np.ndarray:
    """The seeded colour table of an enumerated false-colour subject
    (omv_fc_color_table): size distinct 0xRRGGBB values."""
    out = np.zeros(max(1, size), np.uint32)
    rc = lib().omv_fc_color_table(C.c_uint32(seed), int(subject), C.c_uint32(size), _ptr(out, C.c_uint32))
    if rc != 0:
        raise RuntimeError(f"omv_fc_color_table failed ({rc}): {lib().omv_last_error().decode()}")
    return out[:size]
