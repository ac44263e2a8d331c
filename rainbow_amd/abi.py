"""ctypes mirror of include/rbw.h and a thin loader for librbw.so.

The product path is librbw.so (HIP kernels behind a C ABI).  This module only
declares the structs and signatures; it refuses to run without the built
library, so no Python fallback can stand in for the GPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librbw.so")

RBW_OK = 0
RBW_ERR_INVALID_ARGUMENT = 1
RBW_ERR_OUT_OF_MEMORY = 2
RBW_ERR_DEVICE = 3
RBW_ERR_NO_SCENE = 4
RBW_ERR_UNSUPPORTED = 5
RBW_NONE = 0xFFFFFFFF
RBW_MISS = 0xFFFFFFFE

RBW_SHAPE_NONE, RBW_SHAPE_MESH, RBW_SHAPE_SPHERE = 0, 1, 2
RBW_ENTITY_EMITTER = 1
RBW_ENTITY_REVERSE_ORIENTATION = 2
RBW_ENTITY_HAS_NORMALS = 4
RBW_ENTITY_HAS_UVS = 8
RBW_ENTITY_HAS_MEDIA = 16
RBW_MATERIAL_MATTE, RBW_MATERIAL_METAL, RBW_MATERIAL_PLASTIC = 1, 2, 3
RBW_SAMPLER_PER_SAMPLE_PCG, RBW_SAMPLER_REPLAY, RBW_SAMPLER_STRATIFIED = 0, 1, 2
RBW_TRI_MOLLER_TRUMBORE, RBW_TRI_WATERTIGHT = 0, 1
RBW_SHADOW_CLOSEST_VISIBLE, RBW_SHADOW_ANY_HIT = 0, 1
RBW_INTEGRATOR_PATH, RBW_INTEGRATOR_DIRECT, RBW_INTEGRATOR_VOLUME_PATH = 0, 1, 2
RBW_MEDIUM_HOMOGENEOUS, RBW_MEDIUM_GRID = 0, 1
RBW_KERNEL_NAMES = ["raygen", "extend", "shade", "connect", "film", "guide_infer", "shade_guided", "guide_train"]
RBW_GUIDE_PARAMS = 34160
RBW_KERNEL_COUNT = 8

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class rbw_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("diffuse", C.c_float * 3), ("sigma", C.c_float),
                ("eta", C.c_float * 3), ("k", C.c_float * 3), ("alpha_u", C.c_float),
                ("alpha_v", C.c_float), ("specular", C.c_float * 3), ("ior", C.c_float)]


class rbw_entity(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("first", C.c_uint32), ("count", C.c_uint32),
                ("material", C.c_uint32), ("flags", C.c_uint32), ("radiance", C.c_float * 3)]


class rbw_medium(C.Structure):
    _fields_ = [("type", C.c_uint32), ("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3), ("g", C.c_float),
                ("albedo", C.c_float * 3), ("grid_offset", C.c_uint32), ("grid_size", C.c_uint32 * 3),
                ("world_to_local", C.c_float * 12)]


class rbw_scene_desc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("positions", f32p), ("normals", f32p), ("uvs", f32p),
                ("num_triangles", C.c_uint32), ("indices", u32p), ("triangle_entity", u32p),
                ("triangle_material", u32p), ("num_spheres", C.c_uint32), ("spheres", f32p),
                ("sphere_entity", u32p), ("num_entities", C.c_uint32),
                ("entities", C.POINTER(rbw_entity)), ("num_materials", C.c_uint32),
                ("materials", C.POINTER(rbw_material)), ("num_media", C.c_uint32),
                ("media", C.POINTER(rbw_medium)), ("entity_media", u32p), ("medium_grid_count", C.c_uint32),
                ("medium_grid", f32p)]


class rbw_camera_desc(C.Structure):
    _fields_ = [("raster_to_camera", C.c_float * 16), ("camera_to_world", C.c_float * 16),
                ("width", C.c_uint32), ("height", C.c_uint32), ("bound_min", C.c_int32 * 2),
                ("bound_max", C.c_int32 * 2), ("lens_radius", C.c_float), ("focus_distance", C.c_float)]


class rbw_guiding_opts(C.Structure):
    _fields_ = [("enable", C.c_uint32), ("mis_fraction", C.c_float), ("train", C.c_uint32),
                ("learning_rate", C.c_float), ("batch", C.c_uint32), ("steps_per_pass", C.c_uint32),
                ("scene_min", C.c_float * 3), ("scene_max", C.c_float * 3), ("light_records", C.c_uint32),
                ("guide_depth", C.c_uint32), ("product_targets", C.c_uint32)]


class rbw_render_params(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_threshold", C.c_float),
                ("seed", C.c_uint32), ("sampler", C.c_uint32), ("triangle_test", C.c_uint32),
                ("shadow_mode", C.c_uint32), ("samples_per_pass", C.c_uint32),
                ("replay_offsets", u64p), ("replay_values", f32p), ("replay_count", C.c_uint64),
                ("debug_hits", u32p), ("count_visits", C.c_uint32),
                ("guiding", C.POINTER(rbw_guiding_opts)), ("time_kernels", C.c_uint32), ("bvh_width", C.c_uint32),
                ("strata_x", C.c_uint32), ("strata_y", C.c_uint32), ("dims_1d", C.c_uint32), ("dims_2d", C.c_uint32),
                ("integrator", C.c_uint32), ("emitter_samples", C.c_uint32), ("bsdf_samples", C.c_uint32),
                ("medium_mis", C.c_uint32), ("schedule", C.c_uint32)]


class rbw_stats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("msamples_per_s", C.c_double), ("mrays_per_s", C.c_double),
                ("samples", C.c_uint64), ("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("mis_rays", C.c_uint64), ("light_casts", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64),
                ("extend_node_visits", C.c_uint64), ("extend_prim_tests", C.c_uint64),
                ("kernel_ms", C.c_double * RBW_KERNEL_COUNT),
                ("kernel_launches", C.c_uint64 * RBW_KERNEL_COUNT),
                ("guide_samples", C.c_uint64), ("guide_steps", C.c_uint64), ("guide_loss", C.c_double),
                ("bvh_width", C.c_uint32), ("pad", C.c_uint32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("kernel_ms", "kernel_launches")}
        d["kernel_ms"] = {RBW_KERNEL_NAMES[i]: self.kernel_ms[i] for i in range(RBW_KERNEL_COUNT)}
        d["kernel_launches"] = {RBW_KERNEL_NAMES[i]: self.kernel_launches[i] for i in range(RBW_KERNEL_COUNT)}
        return d


class rbw_device_opts(C.Structure):
    _fields_ = [("device", C.c_int32)]


class rbw_hit(C.Structure):
    _fields_ = [("prim", C.c_uint32), ("entity", C.c_uint32), ("t", C.c_float), ("b1", C.c_float),
                ("b2", C.c_float)]


HIT_DTYPE = np.dtype([("prim", np.uint32), ("entity", np.uint32), ("t", np.float32), ("b1", np.float32),
                      ("b2", np.float32)])


class rbw_job(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_threshold", C.c_float),
                ("seed", C.c_uint32), ("camera", rbw_camera_desc), ("out_rgb", C.POINTER(C.c_float))]


class rbw_batch_opts(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("pool_size", C.c_uint32), ("bvh_width", C.c_uint32),
                ("use_graph", C.c_uint32), ("autotune", C.c_uint32), ("triangle_test", C.c_uint32),
                ("shadow_mode", C.c_uint32), ("guiding", C.POINTER(rbw_guiding_opts))]


class rbw_job_stats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("active_ms", C.c_double), ("msamples_per_s", C.c_double),
                ("first_iteration", C.c_uint32), ("last_iteration", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class rbw_batch_stats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("msamples_per_s", C.c_double), ("mrays_per_s", C.c_double),
                ("autotune_ms", C.c_double), ("samples", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("mis_rays", C.c_uint64), ("iterations", C.c_uint32),
                ("graph_launches", C.c_uint32), ("block_size", C.c_uint32), ("pool_size", C.c_uint32),
                ("bvh_width", C.c_uint32), ("scene_triangles", C.c_uint32), ("guide_samples", C.c_uint64),
                ("guide_steps", C.c_uint32), ("pad", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def ptr(a, t):
    """ctypes pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return C.cast(None, t)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


RBW_STATIC_MAX_RAYS = 262144  # rbw_trace_rays_static's limit per call

# declared C-ABI entry points: name -> (restype, argtypes)
RBW_SIGNATURES = {
    "rbw_create": (C.c_int32, [C.POINTER(rbw_device_opts), C.POINTER(C.c_void_p)]),
    "rbw_destroy": (None, [C.c_void_p]),
    "rbw_last_error": (C.c_char_p, []),
    "rbw_upload_scene": (C.c_int32, [C.c_void_p, C.POINTER(rbw_scene_desc)]),
    "rbw_render": (C.c_int32, [C.c_void_p, C.POINTER(rbw_camera_desc), C.POINTER(rbw_render_params),
                               f32p, C.POINTER(rbw_stats)]),
    "rbw_trace_rays": (C.c_int32, [C.c_void_p, C.c_uint32, f32p, f32p, f32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_void_p]),
    "rbw_trace_rays_static": (C.c_int32, [C.c_void_p, C.c_uint32, f32p, f32p, f32p, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_void_p]),
    "rbw_guide_reset": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "rbw_guide_get_params": (C.c_int32, [C.c_void_p, f32p]),
    "rbw_guide_set_params": (C.c_int32, [C.c_void_p, f32p]),
    "rbw_guide_forward": (C.c_int32, [C.c_void_p, C.c_uint32, f32p, f32p]),
    "rbw_guide_train": (C.c_int32, [C.c_void_p, C.c_uint32, f32p, f32p, f32p, C.c_float, C.c_uint32, f32p, f32p]),
    "rbw_guide_last_ms": (C.c_int32, [C.c_void_p, C.POINTER(C.c_double)]),
    "rbw_guide_jobs_reset": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "rbw_guide_jobs_get_params": (C.c_int32, [C.c_void_p, C.c_uint32, f32p]),
    "rbw_guide_jobs_set_params": (C.c_int32, [C.c_void_p, C.c_uint32, f32p]),
    "rbw_guide_jobs_train": (C.c_int32, [C.c_void_p, C.c_uint32, u32p, f32p, f32p, f32p, C.c_float, C.c_uint32, f32p,
                                         f32p]),
    "rbw_add_scene": (C.c_int32, [C.c_void_p, C.POINTER(rbw_scene_desc), C.POINTER(C.c_uint32)]),
    "rbw_remove_scene": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "rbw_submit_jobs": (C.c_int32, [C.c_void_p, C.POINTER(rbw_job), C.c_uint32, C.POINTER(rbw_batch_opts),
                                    C.POINTER(rbw_job_stats), C.POINTER(rbw_batch_stats)]),
    "rbw_autotune_jobs": (C.c_int32, [C.c_void_p, C.POINTER(rbw_job), C.c_uint32, C.POINTER(rbw_batch_opts),
                                      C.POINTER(C.c_double)]),
    "rbw_sampler_draws": (C.c_int32, [C.c_void_p, C.POINTER(rbw_render_params), C.c_uint32, C.c_uint32, C.c_uint32,
                                      f32p]),
    "rbw_film_to_rgba8": (C.c_int32, [f32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.c_float, C.POINTER(C.c_uint8)]),
    "rbw_write_png": (C.c_int32, [C.c_char_p, f32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.c_float]),
    "rbw_write_pfm": (C.c_int32, [C.c_char_p, f32p, C.c_uint32, C.c_uint32]),
}


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(f"librbw.so is not built ({path}); run __graft_entry__.build() first")
    lib = C.CDLL(path)
    for name, (res, args) in RBW_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
