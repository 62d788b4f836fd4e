"""ctypes bindings of the two native libraries of the package.

* ``libyrh.so``          — the HIP/CDNA4 C ABI declared in ``include/yrh.h``.
* ``libyaoray_host.so``  — the C++20 host library (scene IR, SAH builder,
  workloads) declared in ``yaoray_amd/host/yaoray_host.h``.

Both are built in-tree into ``yaoray_amd/lib`` by ``yaoray_amd.build``. Loading
fails loudly if they are missing: there is no Python or CPU fallback for the
device path.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

YRH_OK = 0
E_INVALID_ARGUMENT = -1
E_HIP = -2
E_UNSUPPORTED = -3
E_NO_SCENE = -4
E_CANCELLED = -5
E_OVERFLOW = -6
YRH_ACCEL_FLAT = 0
YRH_ACCEL_TWO_LEVEL = 1

SCENE_CORNELL = 1
SCENE_TRIANGLES = 2
SCENE_INSTANCED = 3
SCENE_ROOM = 4


class BvhNode(C.Structure):
    _fields_ = [("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3),
                ("payload_offset", C.c_int32), ("primitive_count", C.c_int32)]


class BvhPrimitive(C.Structure):
    _fields_ = [("kind", C.c_int32), ("mesh_primitive", C.c_int32), ("sphere", C.c_int32),
                ("local_triangle", C.c_int32), ("flat_triangle_index", C.c_int32)]


class Vertex(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("normal", C.c_float * 3), ("uv", C.c_float * 2),
                ("tangent", C.c_float * 3), ("tangent_handedness", C.c_float)]


class Primitive(C.Structure):
    _fields_ = [("first_index", C.c_uint32), ("index_count", C.c_uint32), ("material_index", C.c_int32),
                ("has_normals", C.c_uint8), ("has_uvs", C.c_uint8), ("has_tangents", C.c_uint8),
                ("reserved", C.c_uint8)]


class BvhDesc(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_int32),
                ("primitive_indices", C.c_void_p), ("primitive_index_count", C.c_int32),
                ("primitives", C.c_void_p), ("primitive_count", C.c_int32), ("max_depth", C.c_int32)]


class Instance(C.Structure):
    _fields_ = [("primitive", C.c_int32), ("blas_index", C.c_int32), ("object_to_world", C.c_float * 16),
                ("world_to_object", C.c_float * 16), ("world_bounds_min", C.c_float * 3),
                ("world_bounds_max", C.c_float * 3)]


class TlasPrimitive(C.Structure):
    _fields_ = [("kind", C.c_int32), ("instance", C.c_int32), ("sphere", C.c_int32),
                ("world_bounds_min", C.c_float * 3), ("world_bounds_max", C.c_float * 3)]


class GeometryDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_int64),
                ("indices", C.c_void_p), ("index_count", C.c_int64),
                ("primitives", C.c_void_p), ("primitive_count", C.c_int32), ("sphere_count", C.c_int32),
                ("spheres", C.c_void_p), ("accel_kind", C.c_int32), ("flat", BvhDesc),
                ("blases", C.c_void_p), ("blas_count", C.c_int32),
                ("instances", C.c_void_p), ("instance_count", C.c_int32),
                ("tlas", BvhDesc), ("tlas_primitives", C.c_void_p), ("tlas_primitive_count", C.c_int32)]


class Sphere(C.Structure):
    """yrh_sphere == yr::RenderSphere (28 bytes)."""
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("material_index", C.c_int32),
                ("area_light_index", C.c_int32), ("flip_normals", C.c_uint8), ("reserved", C.c_uint8 * 3)]


class RaySoa(C.Structure):
    _fields_ = [("org_tmin", C.c_void_p), ("dir_tmax", C.c_void_p)]


class HitSoa(C.Structure):
    _fields_ = [("hit", C.c_void_p), ("instance", C.c_void_p), ("node_tests", C.c_void_p),
                ("triangle_tests", C.c_void_p), ("sphere_tests", C.c_void_p)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reflectance", C.c_float * 3), ("ior", C.c_float), ("uroughness", C.c_float),
                ("vroughness", C.c_float), ("alpha", C.c_float), ("emission", C.c_float * 3),
                ("absorption_distance", C.c_float), ("absorption_color", C.c_float * 3), ("reserved", C.c_float),
                ("reflectance_texture", C.c_int32), ("alpha_texture", C.c_int32), ("normal_map", C.c_int32),
                ("normal_scale", C.c_float)]


FILTER_NEAREST, FILTER_BILINEAR, FILTER_TRILINEAR, FILTER_EWA = 0, 1, 2, 3
WRAP_REPEAT, WRAP_CLAMP, WRAP_MIRRORED = 0, 1, 2


class Texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("filter", C.c_int32), ("wrap_s", C.c_int32),
                ("wrap_t", C.c_int32), ("reserved", C.c_int32), ("texels", C.c_void_p)]


class Emissive(C.Structure):
    _fields_ = [("primitive_index", C.c_int32), ("radiance", C.c_float * 3), ("area", C.c_float)]


class EnvironmentDesc(C.Structure):
    _fields_ = [("active", C.c_int32), ("radiance", C.c_float * 3), ("strength", C.c_float),
                ("rotation_radians", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("texels", C.c_void_p), ("has_distribution", C.c_int32), ("dist_width", C.c_int32),
                ("dist_height", C.c_int32), ("texel_weights", C.c_void_p), ("row_weights", C.c_void_p),
                ("row_cdf", C.c_void_p), ("conditional_cdfs", C.c_void_p), ("total_weight", C.c_float),
                ("nearest", C.c_int32), ("wrap_s", C.c_int32), ("wrap_t", C.c_int32)]


class AnalyticLight(C.Structure):
    _fields_ = [("kind", C.c_int32), ("position", C.c_float * 3), ("direction", C.c_float * 3),
                ("intensity", C.c_float * 3), ("cone_angle", C.c_float), ("cone_cos_inner", C.c_float)]


class LightTreeNode(C.Structure):
    _fields_ = [("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("power", C.c_float),
                ("left", C.c_int32), ("right", C.c_int32), ("light_index", C.c_int32)]


class ShadingDesc(C.Structure):
    _fields_ = [("materials", C.c_void_p), ("material_count", C.c_int32), ("emissives", C.c_void_p),
                ("emissive_count", C.c_int32), ("alias_probability", C.c_void_p), ("alias_index", C.c_void_p),
                ("alias_pmf", C.c_void_p), ("environment", EnvironmentDesc), ("analytic_lights", C.c_void_p),
                ("analytic_light_count", C.c_int32), ("analytic_alias_probability", C.c_void_p),
                ("analytic_alias_index", C.c_void_p), ("analytic_alias_pmf", C.c_void_p),
                ("analytic_tree", C.c_void_p), ("analytic_tree_count", C.c_int32),
                ("analytic_tree_root", C.c_int32), ("use_analytic_tree", C.c_int32), ("textures", C.c_void_p),
                ("texture_count", C.c_int32)]


class RenderDesc(C.Structure):
    _fields_ = [("camera_origin", C.c_float * 3), ("camera_forward", C.c_float * 3),
                ("camera_right", C.c_float * 3), ("camera_up", C.c_float * 3), ("tan_half_fov", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32), ("sampler", C.c_int32), ("spp", C.c_int32),
                ("max_depth", C.c_int32), ("seed", C.c_uint64), ("radiance_clamp", C.c_float),
                ("sample_index", C.c_int32), ("sample_count", C.c_int32), ("pixel_list", C.c_int32)]


class AdaptiveDesc(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("min_spp", C.c_int32), ("relative_error", C.c_float),
                ("absolute_error", C.c_float), ("confidence", C.c_float)]


class RenderStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("rays_traced", "shadow_rays", "occluded_shadow_rays", "triangle_tests",
                                            "sphere_tests", "bvh_node_tests", "hits", "misses", "bad_samples")]


TRACE_ORDER_INPUT = 0
TRACE_ORDER_ORIGIN = 1
SAMPLER_INDEPENDENT = 0
SAMPLER_STRATIFIED = 1
SAMPLER_OWEN_SOBOL = 2
SAMPLER_ZSOBOL = 3


class SceneInfo(C.Structure):
    _fields_ = [("vertex_count", C.c_int64), ("index_count", C.c_int64), ("primitive_count", C.c_int32),
                ("instance_count", C.c_int32), ("material_count", C.c_int32), ("emissive_count", C.c_int32),
                ("accel_built", C.c_int32), ("accel_kind", C.c_int32), ("node_count", C.c_int32),
                ("max_depth", C.c_int32), ("blas_count", C.c_int32), ("environment_active", C.c_int32),
                ("camera", C.c_float * 13), ("sphere_count", C.c_int32), ("analytic_light_count", C.c_int32),
                ("texture_count", C.c_int32)]


class BatchJob(C.Structure):
    _fields_ = [("camera_origin", C.c_float * 3), ("camera_forward", C.c_float * 3), ("camera_right", C.c_float * 3),
                ("camera_up", C.c_float * 3), ("tan_half_fov", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32), ("spp", C.c_int32)]


class BatchDesc(C.Structure):
    _fields_ = [("jobs", C.c_void_p), ("job_count", C.c_int32), ("sampler", C.c_int32), ("max_depth", C.c_int32),
                ("radiance_clamp", C.c_float), ("seed", C.c_uint64), ("paths_per_pass", C.c_int64),
                ("use_graph", C.c_int32), ("reserved", C.c_int32)]


class GuideDesc(C.Structure):
    _fields_ = [("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("log2_table_size", C.c_int32),
                ("min_resolution", C.c_int32), ("max_resolution", C.c_int32), ("learning_rate", C.c_float),
                ("guide_fraction", C.c_float), ("sample_capacity", C.c_int32), ("seed", C.c_uint64)]


GUIDE_OFF = 0
GUIDE_SAMPLE = 1
GUIDE_RECORD = 2
GUIDE_MLP_PARAMS = 13312


class RenderSettings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("sampler", C.c_int32), ("spp", C.c_int32),
                ("max_depth", C.c_int32), ("radiance_clamp", C.c_float), ("seed", C.c_uint64),
                ("device", C.c_int32), ("requested_backend", C.c_int32), ("adaptive_enabled", C.c_int32),
                ("adaptive_min_spp", C.c_int32), ("adaptive_relative_error", C.c_float),
                ("adaptive_absolute_error", C.c_float), ("adaptive_confidence", C.c_float)]


class RenderRequest(C.Structure):
    _fields_ = [("resume_film", C.c_void_p), ("resume_width", C.c_int32), ("resume_height", C.c_int32),
                ("resume_completed_spp", C.c_int32), ("stop_before_start", C.c_int32),
                ("cancel_after_passes", C.c_int32), ("cancel_error", C.c_char_p)]


class RenderReport(C.Structure):
    _fields_ = ([(k, C.c_int32) for k in ("prepare_ok", "ok", "film_valid", "film_width", "film_height",
                                          "progress_calls", "threads", "bvh_nodes", "bvh_max_depth",
                                          "last_progress_spp")]
                + [(k, C.c_uint64) for k in ("samples_rendered", "rays_traced", "shadow_rays",
                                             "occluded_shadow_rays", "triangle_tests", "sphere_tests",
                                             "bvh_node_tests", "hits", "misses", "bad_samples",
                                             "last_progress_samples", "converged_pixels", "last_progress_active")]
                + [(k, C.c_double) for k in ("elapsed_seconds", "prepare_seconds", "acceleration_build_seconds")]
                + [("error", C.c_char * 256)])


def _sig(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


_VP = C.c_void_p
_I = C.c_int
_I32 = C.c_int32
_I64 = C.c_int64
_U64 = C.c_uint64
_F = C.c_float
_SZ = C.c_size_t

YRH_SIGNATURES = {
    "yrh_create": (_I, [_I, C.POINTER(_VP)]),
    "yrh_destroy": (None, [_VP]),
    "yrh_last_error": (C.c_char_p, [_VP]),
    "yrh_synchronize": (_I, [_VP]),
    "yrh_stream": (_VP, [_VP]),
    "yrh_device_info": (_I, [_VP, C.c_char_p, _I, C.POINTER(_I)]),
    "yrh_device_alloc": (_I, [_VP, _SZ, C.POINTER(_VP)]),
    "yrh_device_free": (_I, [_VP, _VP]),
    "yrh_copy_to_device": (_I, [_VP, _VP, _VP, _SZ]),
    "yrh_copy_to_host": (_I, [_VP, _VP, _VP, _SZ]),
    "yrh_upload_geometry": (_I, [_VP, C.POINTER(GeometryDesc)]),
    "yrh_trace_closest": (_I, [_VP, C.POINTER(RaySoa), _I64, C.POINTER(HitSoa)]),
    "yrh_trace_any": (_I, [_VP, C.POINTER(RaySoa), _I64, _VP, _VP, _VP]),
    "yrh_set_trace_config": (_I, [_VP, _I, _I]),
    "yrh_set_trace_lds_levels": (_I, [_VP, _I]),
    "yrh_set_trace_order": (_I, [_VP, _I]),
    "yrh_set_trace_timing": (_I, [_VP, _I]),
    "yrh_set_render_overlap": (_I, [_VP, _I]),
    "yrh_suggest_paths_per_pass": (_I, [_VP, _I, C.POINTER(C.c_int64)]),
    "yrh_trace_phase_ms": (_I, [_VP, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "yrh_event_create": (_I, [_VP, C.POINTER(_VP)]),
    "yrh_event_destroy": (_I, [_VP, _VP]),
    "yrh_event_record": (_I, [_VP, _VP]),
    "yrh_event_elapsed_ms": (_I, [_VP, _VP, _VP, C.POINTER(_F)]),
    "yrh_upload_shading": (_I, [_VP, C.POINTER(ShadingDesc)]),
    "yrh_film_reset": (_I, [_VP, _I32, _I32]),
    "yrh_film_load": (_I, [_VP, _VP, _I64]),
    "yrh_film_read": (_I, [_VP, _VP, _I64]),
    "yrh_render_pass": (_I, [_VP, C.POINTER(RenderDesc)]),
    "yrh_film_update_active": (_I, [_VP, C.POINTER(AdaptiveDesc), _I32, C.POINTER(_I64), C.POINTER(_I64)]),
    "yrh_render_stats_read": (_I, [_VP, C.POINTER(RenderStats)]),
    "yrh_sample_values": (_I, [_VP, _I32, _U64, _I32, _I64, _VP, _VP, _VP, _VP, _I32, _VP]),
    "yrh_bsdf_query": (_I, [_VP, _VP, _I32, _I64, _VP, _VP, _VP]),
    "yrh_texture_query": (_I, [_VP, _VP, _I64, _VP, _VP, _VP]),
    "yrh_validate_geometry": (_I, [_VP, _VP, _I32]),
    "yrh_trace_closest_noalloc": (_I, [_VP, C.POINTER(RaySoa), _I64, C.POINTER(HitSoa)]),
    "yrh_generate_camera_rays": (_I, [_VP, C.POINTER(RenderDesc), _VP, _VP]),
    "yrh_batch_render": (_I, [_VP, C.POINTER(BatchDesc), _VP, C.POINTER(_I32)]),
    "yrh_batch_film_read": (_I, [_VP, _I32, _VP, _I64]),
    "yrh_guide_configure": (_I, [_VP, C.POINTER(GuideDesc)]),
    "yrh_guide_set_mode": (_I, [_VP, _I32]),
    "yrh_guide_set_cell_budget": (_I, [_VP, _I64]),
    "yrh_guide_set_depth_limit": (_I, [_VP, _I32]),
    "yrh_guide_train": (_I, [_VP, _I32, _I32]),
    "yrh_guide_add_samples": (_I, [_VP, _I64, _VP]),
    "yrh_guide_read_samples": (_I, [_VP, _I64, _VP, C.POINTER(_I64)]),
    "yrh_guide_sample_stats": (_I, [_VP, C.POINTER(_I64), C.POINTER(_F)]),
    "yrh_guide_info": (_I, [_VP, C.POINTER(_I32), C.POINTER(_I64), _VP, _VP]),
    "yrh_guide_params_read": (_I, [_VP, _VP, _VP]),
    "yrh_guide_params_write": (_I, [_VP, _VP, _VP]),
    "yrh_guide_query": (_I, [_VP, _I64, _VP, _VP]),
    "yrh_guide_gradients": (_I, [_VP, _I64, _VP, _F, _VP, _VP, C.POINTER(_F)]),
    "yrh_guide_mixture_eval": (_I, [_VP, _I64, _VP, _VP, _VP, _VP, _VP]),
}

HOST_SIGNATURES = {
    "yrhost_scene_create": (_I, [_I, C.POINTER(_I64), _I, _U64, C.POINTER(_VP)]),
    "yrhost_scene_from_arrays": (_I, [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _VP, _I32, C.POINTER(_VP)]),
    "yrhost_scene_destroy": (None, [_VP]),
    "yrhost_scene_set_camera": (_I, [_VP, _VP]),
    "yrhost_scene_set_materials": (_I, [_VP, _VP, _I32]),
    "yrhost_scene_add_texture": (_I, [_VP, _VP]),
    "yrhost_scene_set_environment_map": (_I, [_VP, _I32, _VP, C.c_float, C.c_float]),
    "yrhost_scene_set_spheres": (_I, [_VP, _VP, _I32]),
    "yrhost_scene_set_analytic_lights": (_I, [_VP, _VP, _I32]),
    "yrhost_scene_spheres": (_I, [_VP, C.POINTER(_VP), C.POINTER(_I32)]),
    "yrhost_scene_set_emissives": (_I, [_VP, _VP, _I32]),
    "yrhost_scene_set_constant_environment": (_I, [_VP, _VP, _F]),
    "yrhost_scene_error": (C.c_char_p, [_VP]),
    "yrhost_scene_info_get": (_I, [_VP, C.POINTER(SceneInfo)]),
    "yrhost_scene_geometry": (_I, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)]),
    "yrhost_build_accel": (_I, [_VP, _I, C.POINTER(C.c_double)]),
    "yrhost_accel_flat": (_I, [_VP, C.POINTER(BvhDesc)]),
    "yrhost_accel_blas": (_I, [_VP, _I32, C.POINTER(BvhDesc)]),
    "yrhost_accel_tlas": (_I, [_VP, C.POINTER(BvhDesc), C.POINTER(_VP), C.POINTER(_I32), C.POINTER(_VP),
                               C.POINTER(_I32)]),
    "yrhost_geometry_desc": (_I, [_VP, C.POINTER(GeometryDesc)]),
    "yrhost_upload_geometry": (_I, [_VP, _VP]),
    "yrhost_shading_desc": (_I, [_VP, C.POINTER(ShadingDesc)]),
    "yrhost_upload_shading": (_I, [_VP, _VP]),
    "yrhost_render_desc": (_I, [_VP, _I32, _I32, _I32, _I32, _I32, _U64, _F, _I32, C.POINTER(RenderDesc)]),
    "yrhost_backend_render": (_I, [_VP, C.POINTER(RenderSettings), C.POINTER(RenderRequest), _VP, _I64,
                                   C.POINTER(RenderReport)]),
    "yrhost_camera_rays": (_I, [_VP, _I, _I, _F, _F, _VP, _VP]),
    "yrhost_random_rays": (_I, [_I64, _U64, _F, _F, _VP, _VP]),
}

_libs: dict[str, C.CDLL] = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def load(name: str, signatures: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    for fn, (res, args) in signatures.items():
        _sig(lib, fn, res, args)
    _libs[name] = lib
    return lib


def yrh() -> C.CDLL:
    return load("libyrh.so", YRH_SIGNATURES)


def host() -> C.CDLL:
    yrh()  # the host library links against libyrh
    return load("libyaoray_host.so", HOST_SIGNATURES)
