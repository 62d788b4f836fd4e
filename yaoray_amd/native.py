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
YRH_ACCEL_FLAT = 0
YRH_ACCEL_TWO_LEVEL = 1

SCENE_CORNELL = 1
SCENE_TRIANGLES = 2
SCENE_INSTANCED = 3


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
                ("accel_kind", C.c_int32), ("flat", BvhDesc),
                ("blases", C.c_void_p), ("blas_count", C.c_int32),
                ("instances", C.c_void_p), ("instance_count", C.c_int32),
                ("tlas", BvhDesc), ("tlas_primitives", C.c_void_p), ("tlas_primitive_count", C.c_int32)]


class RaySoa(C.Structure):
    _fields_ = [("org_tmin", C.c_void_p), ("dir_tmax", C.c_void_p)]


class HitSoa(C.Structure):
    _fields_ = [("hit", C.c_void_p), ("instance", C.c_void_p), ("node_tests", C.c_void_p),
                ("triangle_tests", C.c_void_p)]


class SceneInfo(C.Structure):
    _fields_ = [("vertex_count", C.c_int64), ("index_count", C.c_int64), ("primitive_count", C.c_int32),
                ("instance_count", C.c_int32), ("material_count", C.c_int32), ("emissive_count", C.c_int32),
                ("accel_built", C.c_int32), ("accel_kind", C.c_int32), ("node_count", C.c_int32),
                ("max_depth", C.c_int32), ("blas_count", C.c_int32), ("environment_active", C.c_int32),
                ("camera", C.c_float * 13)]


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
    "yrh_event_create": (_I, [_VP, C.POINTER(_VP)]),
    "yrh_event_destroy": (_I, [_VP, _VP]),
    "yrh_event_record": (_I, [_VP, _VP]),
    "yrh_event_elapsed_ms": (_I, [_VP, _VP, _VP, C.POINTER(_F)]),
}

HOST_SIGNATURES = {
    "yrhost_scene_create": (_I, [_I, C.POINTER(_I64), _I, _U64, C.POINTER(_VP)]),
    "yrhost_scene_from_arrays": (_I, [_VP, _I64, _VP, _I64, _VP, _I32, _VP, _VP, _I32, C.POINTER(_VP)]),
    "yrhost_scene_destroy": (None, [_VP]),
    "yrhost_scene_error": (C.c_char_p, [_VP]),
    "yrhost_scene_info_get": (_I, [_VP, C.POINTER(SceneInfo)]),
    "yrhost_scene_geometry": (_I, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)]),
    "yrhost_build_accel": (_I, [_VP, _I, C.POINTER(C.c_double)]),
    "yrhost_accel_flat": (_I, [_VP, C.POINTER(BvhDesc)]),
    "yrhost_accel_blas": (_I, [_VP, _I32, C.POINTER(BvhDesc)]),
    "yrhost_accel_tlas": (_I, [_VP, C.POINTER(BvhDesc), C.POINTER(_VP), C.POINTER(_I32), C.POINTER(_VP),
                               C.POINTER(_I32)]),
    "yrhost_upload_geometry": (_I, [_VP, _VP]),
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
