"""ctypes binding of libcpt.so (the C ABI declared in include/cpt.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C cudapathtracing_amd/csrc``).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# CPT_LIB_PATH selects another build of the library (e.g. a debug build)
LIB_PATH = os.environ.get("CPT_LIB_PATH") or os.path.join(_HERE, "libcpt.so")

CPT_OK = 0
CPT_ERR_ARG = -1
CPT_ERR_HIP = -2
CPT_ERR_STATE = -3
CPT_ERR_OOM = -4
CPT_ERR_IO = -5
CPT_ERR_ABORTED = -6
CPT_DEVICE_PTRS = 1
NOALLOC_MAX_RAYS = 1 << 20
PERSISTENT_SHADOW_MIN_DEFAULT = 1 << 22

# MaterialType (CudaPathTracing/rtmath.cuh:15) + new entries
M_NONE, M_LIGHT, M_OPAQUE, M_SPECULAR_DIELECTRIC, M_ROUGH_DIELECTRIC, M_VOLUME, M_LAMBERT, M_GGX = range(8)
# ObjectType (CudaPathTracing/Hittable.cuh:8)
OBJ_NONE, OBJ_SPHERE, OBJ_LIGHT, OBJ_TRIANGLE = range(4)

PROFILE_REFERENCE = 0
PROFILE_CONFIG = 1
PROFILE_REAL_TIME = 2
PROFILE_NORMAL = 3
PROFILE_DEPTH = 4
RENDER_COUNT_WORK = 1
RENDER_PACK_AOV = 2

D3 = C.c_double * 3


class Material(C.Structure):
    _fields_ = [("color", D3), ("alpha_x", C.c_double), ("alpha_y", C.c_double),
                ("eta", C.c_double), ("type", C.c_int32), ("pad_", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("center", D3), ("width", C.c_double), ("height", C.c_double), ("normal", D3),
                ("color", D3), ("visible", C.c_int32), ("pad_", C.c_int32)]


class Sphere(C.Structure):
    _fields_ = [("center", D3), ("radius", C.c_double), ("material", C.c_int32), ("pad_", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("background", D3), ("look_from", D3), ("look_at", D3), ("vfov_deg", C.c_double),
                ("aspect", C.c_double), ("width", C.c_int32), ("pad_", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("look_from", D3), ("pixel_center", D3), ("delta_h", D3), ("delta_v", D3),
                ("background", D3), ("width", C.c_int32), ("height", C.c_int32)]


class BuildStats(C.Structure):
    _fields_ = [("morton_ms", C.c_float), ("sort_ms", C.c_float), ("emit_ms", C.c_float),
                ("refit_ms", C.c_float), ("total_ms", C.c_float), ("n_prims", C.c_int32),
                ("cluster_ms", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("rr_depth", C.c_int32), ("profile", C.c_int32), ("rr_prob", C.c_double),
                ("sample_base", C.c_int32), ("want_aov", C.c_int32),
                ("paths_per_batch", C.c_int32), ("flags", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("rays", C.c_double), ("paths", C.c_double),
                ("extend_ms", C.c_double), ("shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("bounces", C.c_int32), ("pad_", C.c_int32), ("extend_rays", C.c_double),
                ("extend_nodes", C.c_double), ("extend_prims", C.c_double),
                ("extend_wave_nodes", C.c_double), ("shadow_rays", C.c_double),
                ("shadow_nodes", C.c_double), ("shadow_prims", C.c_double),
                ("shadow_wave_nodes", C.c_double)]


P = C.c_void_p
dp = C.POINTER(C.c_double)

# (name, restype, argtypes) for every symbol of include/cpt.h
SIGNATURES = [
    ("cpt_create", C.c_int, [C.POINTER(P), C.c_int]),
    ("cpt_destroy", None, [P]),
    ("cpt_last_error", C.c_char_p, [P]),
    ("cpt_set_stream", C.c_int, [P, P]),
    ("cpt_synchronize", C.c_int, [P]),
    ("cpt_version", C.c_char_p, []),
    ("cpt_set_option", C.c_int, [P, C.c_int32, C.c_int32]),
    ("cpt_scene_upload", C.c_int, [P, C.POINTER(Material), C.c_int32, C.POINTER(Light), C.c_int32,
                                   C.POINTER(Sphere), C.c_int32, P, P, C.c_int64]),
    ("cpt_scene_count", C.c_int64, [P]),
    ("cpt_scene_load_ini", C.c_int, [P, C.c_char_p, dp, dp, C.POINTER(CameraDesc)]),
    ("cpt_ini_counts", C.c_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    ("cpt_ini_read", C.c_int, [C.c_char_p, P, P, P, P, P, dp, dp, C.POINTER(CameraDesc)]),
    ("cpt_bvh_build", C.c_int, [P, dp, dp, C.POINTER(BuildStats)]),
    ("cpt_bvh_export", C.c_int, [P, P, P, P, P, P, P]),
    ("cpt_bvh_morton", C.c_int, [P, dp, dp, P, C.c_int]),
    ("cpt_radix_sort_pairs", C.c_int, [P, P, P, C.c_int64, C.c_int32, C.c_int]),
    ("cpt_trace_closest", C.c_int, [P, P, C.c_int64, P, P, C.c_int]),
    ("cpt_trace_count", C.c_int, [P, P, C.c_int64, P, P, C.c_int]),
    ("cpt_last_trace_ms", C.c_float, [P]),
    ("cpt_last_trace_stats", C.c_int, [P, P, P, P]),
    ("cpt_render", C.c_int, [P, C.POINTER(Camera), C.POINTER(RenderParams), P, P,
                             C.POINTER(RenderStats), C.c_int]),
    ("cpt_render_views", C.c_int, [P, C.POINTER(Camera), C.c_int32, C.POINTER(RenderParams), P, P,
                                   C.POINTER(RenderStats), C.c_int]),
    ("cpt_camera_from_desc", C.c_int, [C.POINTER(CameraDesc), C.POINTER(Camera)]),
    ("cpt_tonemap_rgba8", C.c_int, [P, C.c_int32, C.c_int32, P]),
    ("cpt_bilateral", C.c_int, [P, P, P, P, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_int]),
    ("cpt_temporal_blend", C.c_int, [P, P, P, P, P, C.POINTER(Camera), C.POINTER(CameraDesc), C.c_int32, P,
                                     C.c_int]),
    ("cpt_scene_update_triangles", C.c_int, [P, C.c_int64, P, C.c_int64]),
    ("cpt_bvh_refit", C.c_int, [P, C.POINTER(BuildStats)]),
    ("cpt_bvh_cluster_info", C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    ("cpt_denoiser_create", C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(P)]),
    ("cpt_denoiser_destroy", None, [P]),
    ("cpt_denoiser_num_params", C.c_int64, [P]),
    ("cpt_denoiser_get_params", C.c_int, [P, P]),
    ("cpt_denoiser_set_params", C.c_int, [P, P]),
    ("cpt_denoiser_layers", C.c_int, [P, P, P, P, P]),
    ("cpt_denoiser_train_step", C.c_int, [P, P, P, C.c_int32, C.c_int32, C.c_int32, C.c_float, P, P, C.c_int]),
    ("cpt_denoise", C.c_int, [P, P, P, C.c_int32, C.c_int32, C.c_int32, C.c_int]),
    ("cpt_denoiser_stats", C.c_int, [P, P, P, P]),
    ("cpt_conv3x3", C.c_int, [P, C.c_int, P, P, P, P, P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("cpt_gen_icosphere_count", C.c_int64, [C.c_int32]),
    ("cpt_gen_icosphere", C.c_int, [C.c_int32, dp, C.c_double, C.c_double, C.c_uint64, P]),
    ("cpt_gen_triangle_soup", C.c_int, [C.c_int64, C.c_uint64, C.c_double, C.c_double, P]),
    ("cpt_gen_random_rays", C.c_int, [C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_float,
                                      C.c_float, P]),
    ("cpt_tri_bounds", C.c_int, [P, C.c_int64, dp, dp]),
    ("cpt_scene_set_medium", C.c_int, [P, P, C.c_int32, C.c_int32, C.c_int32, dp, dp, C.c_double, dp]),
    ("cpt_gen_fbm_grid", C.c_int, [C.c_int32, C.c_uint64, P]),
    ("cpt_request_abort", C.c_int, [P]),
    ("cpt_clear_abort", C.c_int, [P]),
    ("cpt_write_pfm", C.c_int, [C.c_char_p, P, C.c_int32, C.c_int32]),
    ("cpt_write_png", C.c_int, [C.c_char_p, P, C.c_int32, C.c_int32, C.c_int32]),
    ("cpt_trace_closest_noalloc", C.c_int, [P, P, C.c_int64, P, P]),
]

_lib = None


def lib() -> C.CDLL:
    """Loads libcpt.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(the HIP library has no CPU fallback)")
        # torch (optional) ships its own HIP runtime.  Where torch is
        # installed it is loaded first, so that libcpt.so binds to that same
        # runtime (same soname): two HIP runtimes in one process cannot both
        # open the GPU, and torch tensors are passed to this library as
        # device pointers.  Without torch, libcpt.so loads the system HIP
        # runtime it was linked against.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def d3(v) -> C.Array:
    return D3(*[float(x) for x in v])
