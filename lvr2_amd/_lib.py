"""ctypes binding of liblvrhip.so (declarations: include/lvrhip.h).

The library is required: importing the binding without a built ``liblvrhip.so`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblvrhip.so")

# exported symbols and their ctypes signatures
_SIGS = {
    "lvrhip_last_error": (C.c_char_p, []),
    "lvrhip_create": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "lvrhip_create_device": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "lvrhip_destroy": (C.c_int, [C.c_void_p]),
    "lvrhip_set_k": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "lvrhip_set_flip_point": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "lvrhip_set_method": (C.c_int, [C.c_void_p, C.c_int]),
    "lvrhip_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvrhip_calculate_normals": (C.c_int, [C.c_void_p]),
    "lvrhip_interpolate_normals": (C.c_int, [C.c_void_p]),
    "lvrhip_set_normals": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvrhip_get_normals": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvrhip_get_normals_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvrhip_knn": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p]),
    "lvrhip_distances": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_float, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "lvrhip_distances_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_float, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    "lvrhip_normals_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "lvrhip_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "lvrhip_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "lvrhip_get_timings": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "lvrhip_get_kernel_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "lvrhip_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float]),
    "lvrhip_create_scans": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "lvrhip_create_scans_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                             C.POINTER(C.c_void_p)]),
    "lvrhip_num_scans": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "lvrhip_set_scan_flip_points": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvrhip_rebuild": (C.c_int, [C.c_void_p]),
    "lvrhip_graph_capture": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "lvrhip_graph_launch": (C.c_int, [C.c_void_p]),
    "lvrhip_synchronize": (C.c_int, [C.c_void_p]),
    "lvrhip_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "lvrhip_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "lvrhip_set_retry_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "lvrhip_set_sdf_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
    "lvrhip_get_adaptive_k": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvrhip_grid_build": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "lvrhip_grid_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lvrhip_grid_get_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lvrhip_get_retry_mode": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "lvrhip_distances_backward_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    "lvrhip_radius_search": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_float, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "lvrhip_interpolated_normals": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "lvrhip_set_search_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "lvrhip_grid_build_box": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.c_void_p, C.POINTER(C.c_size_t),
                                        C.POINTER(C.c_size_t)]),
    "lvrhip_mesh_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                    C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "lvrhip_mesh_build_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                         C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "lvrhip_mesh_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "lvrhip_mesh_get_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "lvrhip_distances_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_float, C.c_void_p,
                                                  C.c_void_p, C.c_void_p]),
}

STAT_GROUPS = ("round1", "round2", "round3", "round4", "round5", "merged_rounds", "sdf", "other")
STAT_COUNTERS = ("tiles", "attempts", "resident", "candidates", "refine", "chunks", "resolved", "pops",
                 "cy_trav", "cy_load", "cy_glist", "cy_hist", "cy_select", "cy_collect", "cy_finish", "cy_total")

TIMERS = ("normals_first", "normals_retry", "interpolate", "sdf_kernel", "sdf_sort", "build", "sdf_redo")

EXPORTED = tuple(_SIGS)

_lib = None


class LvrHipError(RuntimeError):
    """Error returned by the C ABI; ``code`` is the LVRHIP_E* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def _bind_process_hip_runtime() -> None:
    """One HIP runtime per process.  torch ships its own libamdhip64.so.7; when torch is
    importable it is imported first so that liblvrhip.so's libamdhip64.so.7 dependency
    resolves to that already-loaded runtime instead of loading a second one."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m lvr2_amd.build`")
        _bind_process_hip_runtime()
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().lvrhip_last_error().decode(errors="replace")
        raise LvrHipError(rc, msg)
