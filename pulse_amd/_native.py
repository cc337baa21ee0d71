"""ctypes binding of libpulse_hip.so (include/pulse_hip.h, include/pulse_scene.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C pulse_amd/csrc``).  Loading fails loudly when it is missing: there is
no CPU fallback for the hot path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpulse_hip.so"

PULSE_OK = 0
ERRORS = {
    -1: "PULSE_E_INVALID_ARG",
    -2: "PULSE_E_HIP",
    -3: "PULSE_E_LAYOUT",
    -4: "PULSE_E_STACK_OVERFLOW",
    -5: "PULSE_E_ITERATION_CAP",
    -6: "PULSE_E_NO_SCENE",
    -7: "PULSE_E_UNSUPPORTED",
    -8: "PULSE_E_BUILD",
}

MODE_REF_GGX = 0
MODE_REF_NEE_LAMBERT = 1
MODE_GGX_NEE = 2
NEE_LAMBERT = 0
NEE_GGX = 1
FLAG_COLLECT_STATS = 1
FLAG_STAGE_TIMING = 2
FLAG_HORIZON_CLAMP = 4
FLAG_ADJOINT_REPARAM = 8
FLAG_LIGHT_POWER = 16
FLAG_COUNT_WORK = 32
FLAG_ADJOINT_ACCUMULATE = 64
FLAG_RECORD_PATHS = 128
PRESENT_RGBA32F = 0
PRESENT_RGBA16F = 1
PRESENT_RGBA8_SRGB = 2
PRESENT_RGBA8_UNORM = 3
GRAD_PER_MATERIAL = 7
OPT_TRAVERSAL_LOOP = 1
OPT_STACK_ENTRIES = 2
OPT_FAST_DIVISION = 3
OPT_PERSISTENT_WAVES = 4
OPT_REFILL_LANES = 5
OPT_GATHER_LANES = 6
OPT_LDS_STACK_ENTRIES = 7
OPT_TRIS_PER_STEP = 8
OPT_ADVANCE_PERIOD = 9
OPT_CONCURRENT_SHADOW = 10
OPT_PAIR_WAVES = 11
OPT_INJECT_ERROR = 12
OPT_PAIR_WAVES_SHADOW = 14
OPT_ADJOINT_REDUCE_ROUNDS = 15
OPT_RECORDED_SAMPLES = 17
OPT_ADJOINT_READ_RECORD = 18
ADJOINT_MAX_VIEWS = 16
ERR_STACK_OVERFLOW = 1
ERR_ITERATION_CAP = 2


class PulseError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class SceneUniform(C.Structure):
    _fields_ = [("instance_count", C.c_uint32), ("light_count", C.c_uint32)]


class BufferView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("bytes", C.c_size_t)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("uniform", SceneUniform),
        ("primitives", BufferView),
        ("triangle_data", BufferView),
        ("triangle_indices", BufferView),
        ("blas_nodes", BufferView),
        ("tlas_nodes", BufferView),
        ("instance_indices", BufferView),
        ("instances", BufferView),
        ("materials", BufferView),
        ("light_emission_strength_cdf", BufferView),
        ("light_triangle_area_cdfs", BufferView),
        ("light_mesh_areas", BufferView),
        ("light_indices", BufferView),
    ]


BUFFER_NAMES = [f[0] for f in SceneDesc._fields_[1:]]


class Camera(C.Structure):
    _fields_ = [
        ("inverse_view_proj", C.c_float * 16),
        ("world_position", C.c_float * 3),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class AdjointView(C.Structure):
    _fields_ = [("camera", Camera), ("image", C.c_void_p), ("target", C.c_void_p), ("replay_rgba", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [
        ("accumulation_count", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("mode", C.c_uint32),
        ("nee_eval", C.c_uint32),
        ("rr_start_depth", C.c_uint32),
        ("flags", C.c_uint32),
        ("_reserved", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_uint64),
        ("extension_rays", C.c_uint64),
        ("incoherent_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("error_flags", C.c_uint32),
        ("stack_entries", C.c_uint32),
        ("ms_total", C.c_double),
        ("ms_primary_extend", C.c_double),
        ("ms_incoherent", C.c_double),
        ("ms_shade", C.c_double),
        ("ms_extend_bounce", C.c_double),
        ("ms_shadow", C.c_double),
        ("extend_launches", C.c_uint32),
        ("shadow_launches", C.c_uint32),
        ("work", (C.c_uint64 * 4) * 3),
    ]


class Material(C.Structure):
    _fields_ = [
        ("base_color", C.c_float * 4),
        ("emissive", C.c_float * 4),
        ("perceptual_roughness", C.c_float),
        ("reflectance", C.c_float),
        ("metallic", C.c_float),
        ("_pad", C.c_float),
    ]


class BuildInfo(C.Structure):
    _fields_ = [
        ("mesh_count", C.c_uint32),
        ("instance_count", C.c_uint32),
        ("light_count", C.c_uint32),
        ("material_count", C.c_uint32),
        ("unique_triangles", C.c_uint64),
        ("instanced_triangles", C.c_uint64),
        ("blas_node_count", C.c_uint32),
        ("tlas_node_count", C.c_uint32),
        ("max_blas_depth", C.c_uint32),
        ("tlas_depth", C.c_uint32),
        ("ms_blas", C.c_double),
        ("ms_instances_tlas", C.c_double),
    ]


# numpy record layouts of the 13 buffers (encase/std430, types.wgsl)
PRIMITIVE_DTYPE = np.dtype([("p_first", "<f4", 3), ("_p0", "<f4"), ("p_second", "<f4", 3), ("_p1", "<f4"),
                            ("p_third", "<f4", 3), ("_p2", "<f4")])
TRIANGLE_DATA_DTYPE = np.dtype([("n_first", "<f4", 3), ("_p0", "<f4"), ("n_second", "<f4", 3), ("_p1", "<f4"),
                                ("n_third", "<f4", 3), ("_p2", "<f4"), ("uv_first", "<f4", 2),
                                ("uv_second", "<f4", 2), ("uv_third", "<f4", 2), ("_p3", "<f4", 2)])
BLAS_NODE_DTYPE = np.dtype([("aabb_min", "<f4", 3), ("a_or_first_tri", "<u4"), ("aabb_max", "<f4", 3),
                            ("tri_count", "<u4")])
TLAS_NODE_DTYPE = np.dtype([("aabb_min", "<f4", 3), ("a_or_first_instance", "<u4"), ("aabb_max", "<f4", 3),
                            ("instance_count", "<u4")])
INSTANCE_DTYPE = np.dtype([("object_world", "<f4", 16), ("world_object", "<f4", 16), ("triangle_offset", "<u4"),
                           ("triangle_count", "<u4"), ("index_offset", "<u4"), ("node_offset", "<u4"),
                           ("material_index", "<u4"), ("_pad", "<u4", 3)])
MATERIAL_DTYPE = np.dtype([("base_color", "<f4", 4), ("emissive", "<f4", 4), ("perceptual_roughness", "<f4"),
                           ("reflectance", "<f4"), ("metallic", "<f4"), ("_pad", "<f4")])
LIGHT_INDEX_DTYPE = np.dtype([("cdf_offset", "<u4"), ("mesh_instance_index", "<u4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("dir", "<f4", 3), ("t_max", "<f4"), ("_pad", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("instance_index", "<u4"), ("triangle_index", "<u4"), ("u", "<f4"),
                      ("v", "<f4")])

BUFFER_DTYPES = {
    "primitives": PRIMITIVE_DTYPE,
    "triangle_data": TRIANGLE_DATA_DTYPE,
    "triangle_indices": np.dtype("<u4"),
    "blas_nodes": BLAS_NODE_DTYPE,
    "tlas_nodes": TLAS_NODE_DTYPE,
    "instance_indices": np.dtype("<u4"),
    "instances": INSTANCE_DTYPE,
    "materials": MATERIAL_DTYPE,
    "light_emission_strength_cdf": np.dtype("<f4"),
    "light_triangle_area_cdfs": np.dtype("<f4"),
    "light_mesh_areas": np.dtype("<f4"),
    "light_indices": LIGHT_INDEX_DTYPE,
}
assert PRIMITIVE_DTYPE.itemsize == 48 and TRIANGLE_DATA_DTYPE.itemsize == 80
assert BLAS_NODE_DTYPE.itemsize == 32 and TLAS_NODE_DTYPE.itemsize == 32
assert INSTANCE_DTYPE.itemsize == 160 and MATERIAL_DTYPE.itemsize == 48
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 20

_P = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int

# (name, restype, argtypes) of every exported symbol declared in include/*.h
class View(C.Structure):
    """pulse_view: one view of a batch."""
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("accum_rgba", C.c_void_p)]


class BatchConfig(C.Structure):
    """pulse_batch_config: launch shape of a batch (does not change images)."""
    _fields_ = [
        ("persistent_waves_per_cu", C.c_uint32),
        ("chunk_rays", C.c_uint32),
        ("stack_entries", C.c_uint32),
        ("traversal_loop", C.c_uint32),
        ("regroup_every", C.c_uint32),
        ("tile_width", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("lds_stack_entries", C.c_uint32),
        ("refill_lanes", C.c_uint32),
        ("_reserved", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_ if not f.startswith("_")}


class GBuffer(C.Structure):
    """pulse_gbuffer: device pointers of width*height float4 arrays."""
    _fields_ = [("position", C.c_void_p), ("normal", C.c_void_p), ("base_color", C.c_void_p), ("emissive", C.c_void_p)]


SIGNATURES = [
    ("pulse_last_error", C.c_char_p, []),
    ("pulse_version", C.c_char_p, []),
    ("pulse_create", _I, [_I, C.POINTER(_P)]),
    ("pulse_destroy", _I, [_P]),
    ("pulse_stream", _P, [_P]),
    ("pulse_sync", _I, [_P]),
    ("pulse_selftest_reciprocal", _I, [_P, C.POINTER(C.c_uint64)]),
    ("pulse_device_alloc", _I, [_P, C.c_size_t, C.POINTER(_P)]),
    ("pulse_device_free", _I, [_P, _P]),
    ("pulse_copy_h2d", _I, [_P, _P, _P, C.c_size_t]),
    ("pulse_copy_d2h", _I, [_P, _P, _P, C.c_size_t]),
    ("pulse_memset_d", _I, [_P, _P, _I, C.c_size_t]),
    ("pulse_scene_upload", _I, [_P, C.POINTER(SceneDesc)]),
    ("pulse_scene_generation", _I, [_P, C.POINTER(C.c_uint64)]),
    ("pulse_scene_update_materials", _I, [_P, _P, _U32]),
    ("pulse_trace", _I, [_P, C.POINTER(Camera), C.POINTER(Params), _P]),
    ("pulse_trace_noalloc", _I, [_P, C.POINTER(Camera), C.POINTER(Params), _P]),
    ("pulse_trace_noalloc_max_pixels", _I, [C.POINTER(C.c_uint64)]),
    ("pulse_trace_rays", _I, [_P, _P, _P, _U32]),
    ("pulse_trace_shadow_rays", _I, [_P, _P, _P, _U32]),
    ("pulse_get_stats", _I, [_P, C.POINTER(Stats)]),
    ("pulse_set_option", _I, [_P, _I, C.c_int64]),
    ("pulse_image_loss", _I, [_P, _P, _P, _U32, C.POINTER(C.c_double)]),
    ("pulse_present", _I, [_P, _P, _U32, _U32, _P, _U32, _U32, _I]),
    ("pulse_trace_gbuffer", _I, [_P, C.POINTER(GBuffer), _U32, _U32, C.POINTER(Params), _P]),
    ("pulse_adjoint", _I, [_P, C.POINTER(Camera), C.POINTER(Params), _P, _P, _P, _P]),
    ("pulse_adjoint_views", _I, [_P, C.POINTER(AdjointView), _U32, C.POINTER(Params), _P]),
    ("pulse_forward_adjoint", _I, [_P, C.POINTER(Camera), C.POINTER(Params), _P, _P, _P, C.POINTER(C.c_double)]),
    ("pulse_get_option", _I, [_P, _I, C.POINTER(C.c_int64)]),
    ("pulse_scene_create", _I, [_P, C.POINTER(SceneDesc), C.POINTER(_P)]),
    ("pulse_scene_destroy", _I, [_P]),
    ("pulse_scene_bind", _I, [_P, _P]),
    ("pulse_scene_update_instances", _I, [_P, _P, C.POINTER(SceneDesc)]),
    ("pulse_scene_refit_instances", _I, [_P, _P, _P, C.c_uint32]),
    ("pulse_scene_rebuild_instances", _I, [_P, _P, _P, C.c_uint32]),
    ("pulse_sb_set_instance_transform", _I, [_P, _U32, _P]),
    ("pulse_sb_update_instances", _I, [_P]),
    ("pulse_batch_default_config", _I, [C.POINTER(BatchConfig)]),
    ("pulse_batch_create", _I, [_P, C.POINTER(View), _U32, C.POINTER(Params), C.POINTER(_P)]),
    ("pulse_batch_destroy", _I, [_P]),
    ("pulse_batch_render", _I, [_P, _U32]),
    ("pulse_batch_set_camera", _I, [_P, _U32, C.POINTER(Camera)]),
    ("pulse_batch_set_config", _I, [_P, C.POINTER(BatchConfig)]),
    ("pulse_batch_get_config", _I, [_P, C.POINTER(BatchConfig)]),
    ("pulse_batch_autotune", _I, [_P, _U32, C.POINTER(BatchConfig), C.POINTER(C.c_double)]),
    ("pulse_batch_get_stats", _I, [_P, C.POINTER(Stats)]),
    ("pulse_trace_batch", _I, [_P, C.POINTER(View), _U32, C.POINTER(Params)]),
    ("pulse_build_blas", _I, [_P, _U32, _P, _U32, C.POINTER(_U32), _P]),
    ("pulse_build_tlas", _I, [_P, _U32, _P, _U32, C.POINTER(_U32), _P]),
    ("pulse_sb_create", _I, [C.POINTER(_P)]),
    ("pulse_sb_destroy", _I, [_P]),
    ("pulse_sb_add_mesh", _I, [_P, _P, _P, _P, _U32, _P, _U32, C.POINTER(_U32)]),
    ("pulse_sb_add_material", _I, [_P, C.POINTER(Material), C.POINTER(_U32)]),
    ("pulse_sb_add_instance", _I, [_P, _U32, _U32, _P]),
    ("pulse_sb_build", _I, [_P, _I]),
    ("pulse_sb_get_desc", _I, [_P, C.POINTER(SceneDesc)]),
    ("pulse_sb_get_info", _I, [_P, C.POINTER(BuildInfo)]),
]

_lib = None


def lib() -> C.CDLL:
    """Loads libpulse_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                              " or `make -C pulse_amd/csrc`")
        L = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != PULSE_OK:
        msg = lib().pulse_last_error()
        raise PulseError(rc, msg.decode() if msg else "")


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
