"""ctypes binding of libmpt.so (include/mpt.h). The library must be built
(``python cadise_amd/build.py``); nothing here falls back to a CPU path."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpt.so")

MPT_OK = 0
MPT_ERR_INVALID, MPT_ERR_PARSE, MPT_ERR_UNSUPPORTED, MPT_ERR_DEVICE, MPT_ERR_IO, MPT_ERR_NOMEM = -1, -2, -3, -4, -5, -6
ERRORS = {-1: "invalid argument", -2: "parse error", -3: "unsupported", -4: "device error", -5: "I/O error",
          -6: "out of memory"}


class MptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Texture(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32),
                ("wrap_clamp", C.c_int32), ("value", C.c_float * 3), ("odd", C.c_void_p), ("even", C.c_void_p),
                ("odd_number", C.c_float), ("even_number", C.c_float)]


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo", Texture), ("f0", C.c_float * 3), ("roughness", Texture),
                ("exponent", C.c_float), ("ratio", C.c_float), ("fresnel", C.c_int32), ("ior_outer", C.c_float),
                ("ior_inner", C.c_float)]


class LightDesc(C.Structure):
    _fields_ = [("prim", C.c_int32), ("color", C.c_float * 3), ("watt", C.c_float), ("back_face", C.c_int32),
                ("type", C.c_int32), ("env", Texture)]


class SceneDesc(C.Structure):
    _fields_ = [("nprims", C.c_int64), ("prim_type", C.c_void_p), ("prim_v", C.c_void_p), ("prim_n", C.c_void_p),
                ("prim_uvw", C.c_void_p), ("prim_mat", C.c_void_p), ("nmats", C.c_int32),
                ("mats", C.POINTER(MaterialDesc)), ("nlights", C.c_int32), ("lights", C.POINTER(LightDesc)),
                ("cluster", C.c_int32), ("cam_pos", C.c_float * 3), ("cam_dir", C.c_float * 3),
                ("cam_up", C.c_float * 3), ("fov_deg", C.c_float), ("sensor_width_mm", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32), ("tile_x", C.c_int32), ("tile_y", C.c_int32),
                ("filter", C.c_int32), ("filter_w", C.c_float), ("filter_h", C.c_float), ("sigma_x", C.c_float),
                ("sigma_y", C.c_float), ("amplitude", C.c_float), ("mitchell_b", C.c_float),
                ("mitchell_c", C.c_float), ("sampler", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("split_sah", C.c_int32), ("estimator", C.c_int32), ("output_filename", C.c_char_p)]


class BvhOpts(C.Structure):
    _fields_ = [("width", C.c_int32), ("leaf_max", C.c_int32), ("block", C.c_int32), ("stack_lds", C.c_int32),
                ("blocks_per_cu", C.c_int32)]


class SceneStats(C.Structure):
    _fields_ = [("nprims", C.c_int64), ("nlights", C.c_int64), ("binary_nodes", C.c_int64),
                ("wide_nodes", C.c_int64), ("binary_depth", C.c_int32), ("wide_depth", C.c_int32),
                ("width", C.c_int32), ("leaf_max", C.c_int32), ("build_ms", C.c_double), ("collapse_ms", C.c_double),
                ("device_bytes", C.c_int64), ("image_width", C.c_int32), ("image_height", C.c_int32),
                ("spp", C.c_int32), ("max_depth", C.c_int32), ("ntextures", C.c_int64),
                ("tex_elements", C.c_int64), ("stack_lds", C.c_int32), ("trace_block", C.c_int32),
                ("blocks_per_cu", C.c_int32)]


class RenderParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("spp_begin", C.c_int32), ("spp_end", C.c_int32), ("max_depth", C.c_int32),
                ("accumulate", C.c_int32), ("flags", C.c_int32)]


MPT_RENDER_KEEP_SAMPLES = 1
MPT_RENDER_PASSWISE = 2


class Stats(C.Structure):
    _fields_ = [("camera_rays", C.c_int64), ("continuation_rays", C.c_int64), ("shadow_rays", C.c_int64),
                ("mis_rays", C.c_int64), ("samples", C.c_int64), ("box_tests", C.c_int64), ("tri_tests", C.c_int64),
                ("box_tests_any", C.c_int64), ("tri_tests_any", C.c_int64), ("closest_launches", C.c_int64),
                ("anyhit_launches", C.c_int64),
                ("ms_total", C.c_double), ("ms_closest", C.c_double), ("ms_anyhit", C.c_double),
                ("ms_shade", C.c_double), ("ms_camera", C.c_double), ("ms_film", C.c_double),
                ("ms_adjoint", C.c_double), ("ms_gpu", C.c_double), ("path_slots", C.c_int64),
                ("mis_resolved", C.c_int64), ("shadow_resolved", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOpts(C.Structure):
    _fields_ = [("width", C.c_int32), ("stack_lds", C.c_int32), ("block", C.c_int32), ("capacity_log2", C.c_int32),
                ("graph_steps", C.c_int32), ("blocks_per_cu", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchTrial(C.Structure):
    _fields_ = [("opts", BatchOpts), ("ms", C.c_double)]


EXPORTS = ["mpt_create", "mpt_destroy", "mpt_last_error", "mpt_version", "mpt_scene_load_crsd", "mpt_scene_parse_crsd",
           "mpt_scene_create", "mpt_scene_destroy", "mpt_scene_stats_get", "mpt_scene_binary_bvh", "mpt_scene_camera",
           "mpt_scene_light_radiance", "mpt_scene_light_pick_pdf", "mpt_trace", "mpt_intersect_detail",
           "mpt_trace_device", "mpt_trace_device_noalloc", "mpt_render", "mpt_film_device", "mpt_film_srgb8", "mpt_film_srgb8_device", "mpt_film_save",
           "mpt_scene_output_filename", "mpt_image_write", "mpt_stats_get", "mpt_release_stream_buffers",
           "mpt_scene_texture_info", "mpt_scene_set_texture", "mpt_scene_get_texture", "mpt_scene_set_light_radiance",
           "mpt_render_adjoint",
           "mpt_batch_create", "mpt_batch_destroy", "mpt_batch_render", "mpt_batch_film", "mpt_batch_autotune",
           "mpt_batch_get_opts", "mpt_batch_set_opts"]

_lib = None


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python cadise_amd/build.py` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.mpt_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.mpt_destroy.argtypes = [vp]
    L.mpt_last_error.restype = C.c_char_p
    L.mpt_last_error.argtypes = [vp]
    L.mpt_version.restype = C.c_char_p
    L.mpt_scene_load_crsd.argtypes = [vp, C.c_char_p, C.POINTER(BvhOpts), C.POINTER(vp)]
    L.mpt_scene_parse_crsd.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(BvhOpts), C.POINTER(vp)]
    L.mpt_scene_create.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(BvhOpts), C.POINTER(vp)]
    L.mpt_scene_destroy.argtypes = [vp]
    L.mpt_scene_stats_get.argtypes = [vp, C.POINTER(SceneStats)]
    L.mpt_scene_binary_bvh.argtypes = [vp, C.POINTER(i64), vp, vp, vp, vp, vp]
    L.mpt_scene_camera.argtypes = [vp, vp, vp]
    L.mpt_scene_light_radiance.argtypes = [vp, vp]
    L.mpt_scene_light_pick_pdf.argtypes = [vp, vp]
    L.mpt_scene_get_texture.argtypes = [vp, i32, vp]
    L.mpt_trace.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
    L.mpt_trace_device.argtypes = [vp, vp, C.c_size_t, C.c_int, vp, vp]
    L.mpt_trace_device_noalloc.argtypes = [vp, vp, C.c_size_t, C.c_int, vp, vp]
    L.mpt_intersect_detail.argtypes = [vp, vp, C.c_size_t, vp]
    L.mpt_render.argtypes = [vp, C.POINTER(RenderParams), vp]
    L.mpt_film_device.argtypes = [vp, C.POINTER(vp)]
    L.mpt_film_srgb8.argtypes = [vp, vp]
    L.mpt_film_save.argtypes = [vp, C.c_char_p]
    L.mpt_film_srgb8_device.argtypes = [vp, vp, i64, vp, vp]
    L.mpt_scene_output_filename.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.mpt_image_write.argtypes = [C.c_char_p, vp, i32, i32]
    L.mpt_stats_get.argtypes = [vp, C.POINTER(Stats)]
    L.mpt_release_stream_buffers.argtypes = [vp]
    L.mpt_scene_texture_info.argtypes = [vp, i32, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    L.mpt_scene_set_texture.argtypes = [vp, i32, vp]
    L.mpt_scene_set_light_radiance.argtypes = [vp, i32, vp]
    L.mpt_render_adjoint.argtypes = [vp, C.POINTER(RenderParams), vp, vp, vp]
    L.mpt_batch_create.argtypes = [vp, C.POINTER(vp), i32, C.POINTER(BatchOpts), C.POINTER(vp)]
    L.mpt_batch_destroy.argtypes = [vp]
    L.mpt_batch_render.argtypes = [vp, vp, i32, C.POINTER(Stats)]
    L.mpt_batch_film.argtypes = [vp, i32, vp]
    L.mpt_batch_autotune.argtypes = [vp, vp, i32, C.POINTER(BatchOpts), C.POINTER(BatchTrial), i32, C.POINTER(i32)]
    L.mpt_batch_get_opts.argtypes = [vp, C.POINTER(BatchOpts)]
    L.mpt_batch_set_opts.argtypes = [vp, C.POINTER(BatchOpts)]
    _lib = L
    return L
