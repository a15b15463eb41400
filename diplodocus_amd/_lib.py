"""ctypes binding of libdpl.so (include/dpl/dpl.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C diplodocus_amd/csrc``).
Loading fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdpl.so")

DPL_OK, DPL_ERR_INVALID_ARG, DPL_ERR_OOM, DPL_ERR_DEVICE, DPL_ERR_UNSUPPORTED, DPL_ERR_INTERNAL = range(6)
DPL_MEM_HOST, DPL_MEM_DEVICE = 0, 1
DPL_TRACE_ANY, DPL_TRACE_CULL_BACK, DPL_TRACE_WIDE, DPL_TRACE_COHERENT = 1, 2, 4, 8

STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "DEVICE", 4: "UNSUPPORTED", 5: "INTERNAL"}


class BuildParams(C.Structure):
    _fields_ = [("morton_type", C.c_int32), ("bv_type", C.c_int32), ("nn_search_radius", C.c_int32),
                ("max_triangles_per_leaf", C.c_int32), ("kdop_size", C.c_int32), ("block_size", C.c_int32),
                ("sort_digit_bits", C.c_int32), ("keep_intermediates", C.c_int32)]


class BuildStats(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("upload_ms", C.c_double), ("stage_ms", C.c_double * 8),
                ("node_count", C.c_int32), ("inner_node_count", C.c_int32), ("leaf_node_count", C.c_int32),
                ("ploc_iterations", C.c_int32), ("memory_consumption", C.c_int64), ("sah_cost", C.c_double),
                ("tree_height", C.c_int32), ("ploc_launches", C.c_int32), ("ploc_launch_ms", C.c_double),
                ("ploc_bytes", C.c_double)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "stage_ms"}
        d["stage_ms"] = dict(zip(["bounds", "codes", "sort", "ploc", "collapse", "layout", "sobb", "stats"],
                                 list(self.stage_ms)))
        return d


class BvhLayout(C.Structure):
    _fields_ = [("prim_count", C.c_int64), ("node_count", C.c_int64), ("ploc_node_count", C.c_int64),
                ("iteration_count", C.c_int64), ("mesh_count", C.c_int32), ("node_bytes", C.c_int32)]


class WideInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("node_bytes", C.c_int32), ("node_count", C.c_int64), ("widen_ms", C.c_double)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("seed", C.c_uint64), ("camera", C.c_float * 11), ("material_count", C.c_int32),
                ("reserved", C.c_int32)]


class WhittedParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("max_depth", C.c_int32),
                ("area_light_sample_cnt", C.c_int32), ("backface_culling", C.c_int32), ("reserved", C.c_int32 * 3),
                ("background", C.c_float * 3), ("camera", C.c_float * 11)]


class WhittedStats(C.Structure):
    _fields_ = [("primary_ray_count", C.c_int64), ("secondary_ray_count", C.c_int64),
                ("shadow_ray_count", C.c_int64), ("query_count", C.c_int64), ("traversal_count", C.c_int64),
                ("intersection_count", C.c_int64), ("raytracing_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshInfo(C.Structure):
    _fields_ = [("node_offset", C.c_int64), ("prim_offset", C.c_int64), ("node_count", C.c_int32),
                ("prim_count", C.c_int32), ("tree_height", C.c_int32), ("ploc_iterations", C.c_int32),
                ("sah_cost", C.c_double)]


class AutotuneInfo(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("sah_cost", C.c_double), ("candidates", C.c_int32),
                ("from_cache", C.c_int32)]


class DplError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"dpl error {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C diplodocus_amd/csrc); there is no fallback path")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname libamdhip64.so.7,
    # but its libtorch_hip asks for the unversioned name). Loading torch first makes libdpl's
    # libamdhip64.so.7 dependency resolve to the already-loaded copy; loading libdpl first
    # would pull in a second runtime and torch would then see no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, U32, U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
    sig = {
        "dpl_build_params_from_type": (I32, [I32, I32, I32, I32, C.POINTER(BuildParams)]),
        "dpl_context_create": (I32, [I32, C.POINTER(P)]),
        "dpl_context_destroy": (None, [P]),
        "dpl_last_error": (C.c_char_p, [P]),
        "dpl_context_set_stream": (I32, [P, P]),
        "dpl_context_synchronize": (I32, [P]),
        "dpl_context_cu_count": (I32, [P]),
        "dpl_bvh_build": (I32, [P, P, I64, I32, C.POINTER(BuildParams), C.POINTER(P), C.POINTER(BuildStats)]),
        "dpl_bvh_rebuild": (I32, [P, P, I64, I32, C.POINTER(BuildParams), C.POINTER(BuildStats)]),
        "dpl_bvh_build_batched": (I32, [P, P, P, I32, I32, C.POINTER(BuildParams), C.POINTER(P),
                                        C.POINTER(BuildStats)]),
        "dpl_bvh_destroy": (None, [P]),
        "dpl_bvh_get_stats": (I32, [P, C.POINTER(BuildStats)]),
        "dpl_bvh_sah_aabb": (I32, [P, C.POINTER(C.c_double)]),
        "dpl_bvh_get_layout": (I32, [P, C.POINTER(BvhLayout)]),
        "dpl_bvh_download": (I32, [P, P, P, P, P, P, P, P]),
        "dpl_trace": (I32, [P, P, P, I64, U32, P, I32]),
        "dpl_trace_noalloc": (I32, [P, P, P, I64, U32, P]),
        "dpl_trace_counters": (I32, [P, P, P, I64, U32, P, I32, P, P]),
        "dpl_trace_batched": (I32, [P, P, P, P, I64, U32, P, I32]),
        "dpl_compute_codes": (I32, [P, P, I64, I32, P, P]),
        "dpl_radix_sort_pairs": (I32, [P, P, P, I64, I32, I32]),
        "dpl_render_forward": (I32, [P, P, C.POINTER(RenderParams), P, P, P, P]),
        "dpl_render_adjoint": (I32, [P, P, C.POINTER(RenderParams), P, P, P, P, P, P]),
        "dpl_bvh_get_meshes": (I32, [P, P]),
        "dpl_render_whitted": (I32, [P, P, C.POINTER(WhittedParams), P, P, P, P, I32, P, I32, P]),
        "dpl_render_whitted_stats": (I32, [P, P, C.POINTER(WhittedParams), P, P, P, P, I32, P, I32, P,
                                           C.POINTER(WhittedStats)]),
        "dpl_trace_graph_create": (I32, [P, P, P, P, I64, U32, P, C.POINTER(P)]),
        "dpl_trace_graph_launch": (I32, [P]),
        "dpl_bvh_widen": (I32, [P, I32, C.POINTER(WideInfo)]),
        "dpl_bvh_get_wide": (I32, [P, C.POINTER(WideInfo)]),
        "dpl_bvh_download_wide": (I32, [P, P]),
        "dpl_trace_graph_destroy": (None, [P]),
        "dpl_autotune": (I32, [P, P, I64, P, I32, I32, C.POINTER(BuildParams), C.c_double, C.c_char_p,
                               C.POINTER(BuildParams), C.POINTER(AutotuneInfo)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Every symbol the public header declares (checked by the CPU test suite).
EXPORTED = [
    "dpl_build_params_from_type", "dpl_context_create", "dpl_context_destroy", "dpl_last_error",
    "dpl_context_set_stream", "dpl_context_synchronize", "dpl_context_cu_count", "dpl_bvh_build", "dpl_bvh_rebuild",
    "dpl_bvh_build_batched", "dpl_bvh_destroy", "dpl_bvh_get_stats", "dpl_trace", "dpl_trace_counters",
    "dpl_trace_batched",
    "dpl_bvh_get_layout", "dpl_bvh_download", "dpl_compute_codes", "dpl_radix_sort_pairs", "dpl_render_forward",
    "dpl_render_adjoint", "dpl_bvh_get_meshes", "dpl_trace_graph_create", "dpl_trace_graph_launch",
    "dpl_trace_graph_destroy", "dpl_autotune", "dpl_render_whitted", "dpl_render_whitted_stats", "dpl_bvh_widen",
    "dpl_bvh_get_wide", "dpl_bvh_download_wide", "dpl_bvh_sah_aabb", "dpl_trace_noalloc",
]
