"""ctypes binding of libsprout_gpu.so (the C ABI declared in include/sprout_gpu.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsprout_gpu.so")

MI_NULL = 0xFFFFFFFF

STATUS = {
    0: "MI_OK",
    -1: "MI_ERR_NO_CONTEXT",
    -2: "MI_ERR_ARGS",
    -3: "MI_ERR_HIP",
    -4: "MI_ERR_HANDLE",
    -5: "MI_ERR_UNSPLITTABLE",
    -6: "MI_ERR_STACK",
}


class MiError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)}: {message}")


class MiRay(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("min_t", ctypes.c_float),
                ("direction", ctypes.c_float * 3), ("max_t", ctypes.c_float)]


class MiHit(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("u", ctypes.c_float), ("v", ctypes.c_float),
                ("triangle", ctypes.c_uint32)]


class MiBvhSettings(ctypes.Structure):
    _fields_ = [("num_slices", ctypes.c_uint32), ("sweep_threshold", ctypes.c_uint32),
                ("max_primitives", ctypes.c_uint32)]


class MiBvhInfo(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint32), ("num_slots", ctypes.c_uint32),
                ("num_triangles", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("aabb_min", ctypes.c_float * 3), ("aabb_max", ctypes.c_float * 3),
                ("build_ms", ctypes.c_double)]


class MiTraceConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("kernel", ctypes.c_uint32), ("refill_lanes", ctypes.c_uint32),
                ("claim_batch", ctypes.c_uint32), ("lds_entries", ctypes.c_uint32),
                ("sort_rays", ctypes.c_uint32)]


class MiMaterial(ctypes.Structure):
    _fields_ = [("color", ctypes.c_float * 3), ("emission", ctypes.c_float * 3),
                ("roughness", ctypes.c_float), ("metallic", ctypes.c_float),
                ("ior", ctypes.c_float), ("kind", ctypes.c_uint32)]


class MiMesh(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("num_vertices", ctypes.c_uint32),
                ("indices", ctypes.c_void_p), ("num_triangles", ctypes.c_uint32)]


class MiSensor(ctypes.Structure):
    _fields_ = [("filter", ctypes.c_uint32), ("radius", ctypes.c_float), ("alpha", ctypes.c_float),
                ("b", ctypes.c_float), ("c", ctypes.c_float), ("clamp_max", ctypes.c_float * 3)]


MI_FILTER_BOX = 0
MI_FILTER_GAUSSIAN = 1
MI_FILTER_BLACKMAN = 2
MI_FILTER_MITCHELL = 3


class MiCamera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("position", ctypes.c_float * 3), ("rotation", ctypes.c_float * 9),
                ("fov_deg", ctypes.c_float)]


class MiRenderSettings(ctypes.Structure):
    _fields_ = [("samples_per_pixel", ctypes.c_uint32), ("samples_per_wave", ctypes.c_uint32),
                ("first_sample", ctypes.c_uint32), ("min_bounces", ctypes.c_uint32),
                ("max_bounces", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class MiRenderStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("paths", ctypes.c_double),
                ("extend_rays", ctypes.c_double), ("shadow_rays", ctypes.c_double),
                ("trained", ctypes.c_double)]


class MiGuideSettings(ctypes.Structure):
    _fields_ = [("guide", ctypes.c_uint32), ("fraction", ctypes.c_float),
                ("train", ctypes.c_uint32), ("train_steps", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("learning_rate", ctypes.c_float),
                ("record_depth", ctypes.c_uint32), ("guide_depth", ctypes.c_uint32),
                ("medium", ctypes.c_uint32)]


MI_MATERIAL_SUBSTITUTE = 0
MI_MATERIAL_LAMBERT = 1
MI_MATERIAL_LIGHT = 2
MI_RENDER_NO_GRAPH = 1
MI_RENDER_NO_LIGHT_SPLIT = 2
MI_RENDER_GOLDEN_RATIO = 4
MI_RENDER_REGENERATE = 8
MI_RENDER_SPECTRAL = 16

_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_vp = ctypes.c_void_p
_pu32 = ctypes.POINTER(ctypes.c_uint32)

# name -> (restype, argtypes); every function of include/sprout_gpu.h
SIGNATURES = {
    "mi_init": (_i32, [_i32]),
    "mi_release": (_i32, []),
    "mi_last_error": (ctypes.c_char_p, []),
    "mi_bvh_build": (_i32, [_vp, _u32, _u32, _vp, _u32, ctypes.POINTER(MiBvhSettings), _vp, _pu32]),
    "mi_bvh_upload": (_i32, [_vp, _u32, _vp, _u32, _vp, _u32, _u32, _vp, _u32, _pu32]),
    "mi_bvh_release": (_i32, [_u32]),
    "mi_bvh_get_info": (_i32, [_u32, ctypes.POINTER(MiBvhInfo)]),
    "mi_bvh_download": (_i32, [_u32, _vp, _vp]),
    "mi_trace_closest": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_trace_closest_noalloc": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_trace_any": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_trace_counts": (_i32, [_u32, _vp, _u32, _i32, _vp, _vp, _vp]),
    "mi_set_trace_config": (_i32, [ctypes.POINTER(MiTraceConfig)]),
    "mi_trace_autotune": (_i32, [_u32, _vp, _u32, _u32, _vp, ctypes.POINTER(MiTraceConfig),
                                 ctypes.POINTER(ctypes.c_float)]),
    "mi_sync": (_i32, [_vp]),
    "mi_gen_displaced_icosphere": (_i32, [_u32, _u32, _f32, _u64, _u32, _f32, _f32, _u64, _vp,
                                          _pu32, _vp, _pu32]),
    "mi_gen_camera_rays": (_i32, [_u32, _u32, _vp, _vp, _vp, _f32, _vp]),
    "mi_gen_bounce_rays": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _u64, _vp, _u32]),
    "mi_scene_create": (_i32, [_vp, _u32, _vp, _u32, _vp, _vp, _u32,
                               ctypes.POINTER(MiBvhSettings), _pu32]),
    "mi_scene_create_instanced": (_i32, [ctypes.POINTER(MiMesh), _u32, _vp, _vp, _vp, _vp, _u32,
                                         _vp, _u32, _vp, _u32, _vp, _vp, _u32,
                                         ctypes.POINTER(MiBvhSettings), _pu32]),
    "mi_scene_update_instances": (_i32, [_u32, _vp, _vp, _u32]),
    "mi_scene_release": (_i32, [_u32]),
    "mi_render_filtered": (_i32, [_u32, ctypes.POINTER(MiCamera), ctypes.POINTER(MiRenderSettings),
                                  ctypes.POINTER(MiSensor), _vp, _vp, ctypes.POINTER(MiRenderStats)]),
    "mi_render": (_i32, [_u32, ctypes.POINTER(MiCamera), ctypes.POINTER(MiRenderSettings), _vp,
                         _vp, ctypes.POINTER(MiRenderStats)]),
    "mi_render_guided": (_i32, [_u32, ctypes.POINTER(MiCamera), ctypes.POINTER(MiRenderSettings),
                                ctypes.POINTER(MiGuideSettings), _vp, _vp,
                                ctypes.POINTER(MiRenderStats)]),
    "mi_light_tree_build": (_i32, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _pu32]),
    "mi_light_part_tree_build": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _pu32, _pu32]),
    "mi_light_tree_build_device": (_i32, [_vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp,
                                          _pu32, _vp]),
    "mi_light_part_trees_build_device": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp,
                                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mi_scene_update_positions": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_light_tree_sample": (_i32, [_u32, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp]),
    "mi_light_tree_pdf": (_i32, [_u32, _vp, _vp, _vp, _u32, _u32, _vp, _vp]),
    "mi_scene_light_tree": (_i32, [_u32, _vp, _vp, _vp, _vp, _vp, _pu32, _pu32]),
    "mi_instances_create": (_i32, [_vp, _vp, _vp, _u32, ctypes.POINTER(MiBvhSettings), _vp, _pu32]),
    "mi_instances_release": (_i32, [_u32]),
    "mi_trace_instances_closest": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_trace_instances_any": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_instances_download": (_i32, [_u32, _vp, _vp, _vp, _pu32, _pu32]),
    "mi_scene_set_normals": (_i32, [_u32, _vp, _u32]),
    "mi_scene_set_mesh_normals": (_i32, [_u32, _u32, _vp, _u32]),
    "mi_scene_set_medium": (_i32, [_u32, _vp, _vp, _vp, _vp, _vp, _vp, _f32]),
    "mi_medium_gridtree": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _pu32, _vp, _pu32]),
    "mi_medium_majorant_tree": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _pu32, _vp, _pu32]),
    "mi_scene_set_medium_majorants": (_i32, [_u32, _vp, _vp, _vp, _vp, _vp, _vp, _f32, _u32]),
    "mi_medium_track": (_i32, [_u32, _vp, _u32, ctypes.c_uint64, _u32, _u32, _vp, _vp]),
    "mi_ggx_e_table": (_i32, [_vp, _vp]),
    "mi_substitute_evaluate": (_i32, [_u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mi_substitute_sample": (_i32, [_u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mi_gen_room": (_i32, [_f32, _vp, _pu32, _vp, _pu32, _vp, _vp, _pu32]),
    "mi_gen_cloud_scene": (_i32, [_vp, _pu32, _vp, _pu32, _vp, _vp, _pu32]),
    "mi_gen_fbm_density": (_i32, [_u32, _u32, _u32, _u32, ctypes.c_uint64, _f32, _vp]),
    "mi_guide_create": (_i32, [_u64, _pu32]),
    "mi_guide_release": (_i32, [_u32]),
    "mi_guide_set_params": (_i32, [_u32, _vp]),
    "mi_guide_get_params": (_i32, [_u32, _vp]),
    "mi_guide_set_frequencies": (_i32, [_u32, _vp]),
    "mi_guide_forward": (_i32, [_u32, _vp, _u32, _vp, _vp]),
    "mi_guide_gradients": (_i32, [_u32, _vp, _vp, _vp, _u32, _vp, _vp]),
    "mi_guide_train_step": (_i32, [_u32, _vp, _vp, _vp, _u32, _f32, _vp]),
}

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m sprout_amd.build` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int, where: str) -> None:
    if status != 0:
        msg = load().mi_last_error()
        raise MiError(status, where, msg.decode() if msg else "")
