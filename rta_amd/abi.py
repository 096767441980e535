"""ctypes bindings of the bbvh-hip C-ABI (include/rta_hip.h, include/rta_host.h).

The shared libraries are built in-tree into rta_amd/lib by
`make -C rta_amd/csrc`.  Loading fails loudly when they are missing: there is
no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")
HIP_LIB_PATH = os.path.join(LIB_DIR, "librta_hip.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "librta_host.so")

# --------------------------------------------------------------- layouts
# byte images of the C structs (librta/basic_types.h:67-95,
# bbvh/bbvh-node.h:5-43, librta/cuda-vec.h:30-43)
TRI_DTYPE = np.dtype(
    [("a", "<f4", 3), ("b", "<f4", 3), ("c", "<f4", 3),
     ("na", "<f4", 3), ("nb", "<f4", 3), ("nc", "<f4", 3),
     ("ta", "<f4", 2), ("tb", "<f4", 2), ("tc", "<f4", 2),
     ("material_index", "<i4"), ("_pad", "<i4")])
AABB_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3)])
NODE_DTYPE = np.dtype([("type_left_elems", "<u4"), ("right_tris", "<u4"),
                       ("min", "<f4", 3), ("max", "<f4", 3)])
HIT_DTYPE = np.dtype([("t", "<f4"), ("beta", "<f4"), ("gamma", "<f4"), ("ref", "<u4")])
# rta_treelet (include/rta_hip.h): root box, root pair-record word, first position
TREELET_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("word", "<u4"), ("first", "<u4")])
TIMES_FIELDS = ("init_triangles", "voxel_bounds", "morton", "sort", "tree", "boxes",
                "permute", "total")
assert TRI_DTYPE.itemsize == 104 and AABB_DTYPE.itemsize == 24
assert NODE_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16
assert TREELET_DTYPE.itemsize == 32

FLT_MAX = np.float32(np.finfo(np.float32).max)

RTA_OK = 0
RTA_E_INVALID = -1
RTA_E_WORKSPACE = -2
RTA_E_NOT_READY = -3
RTA_E_STACK = -4
RTA_LBVH_CODE_ZERO = 1
RTA_LBVH_FRONTIER_OVERFLOW = 2
RTA_LBVH_NOALLOC_MAX_TRIS = 262144


class RtaTimes(C.Structure):
    _fields_ = [(f, C.c_float) for f in TIMES_FIELDS]

    def as_dict(self):
        return {f: float(getattr(self, f)) for f in TIMES_FIELDS}


class TraceConfig(C.Structure):
    """rta_trace_config: launch shape of the persistent tracer (stack 0 =
    the full 32 entries; 16 or 24 for batch traversal)."""
    _fields_ = [("block", C.c_int), ("refill_min", C.c_int), ("steps", C.c_int),
                ("blocks_per_cu", C.c_int), ("leaf_min", C.c_int), ("stack", C.c_int)]

    def as_tuple(self):
        return (self.block, self.refill_min, self.steps, self.blocks_per_cu, self.leaf_min,
                self.stack)


class RtaError(RuntimeError):
    """A non-zero status from the C-ABI (negative: argument/state, positive: HIP)."""

    def __init__(self, code: int, func: str, message: str):
        super().__init__(f"{func} failed with status {code}: {message}")
        self.code = code
        self.func = func


_vp, _sz, _i, _u32, _u64, _f = C.c_void_p, C.c_size_t, C.c_int, C.c_uint32, C.c_uint64, C.c_float
_pi, _psz, _pf = C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_float)

# name -> argtypes (every function returns int status)
HIP_SIGNATURES = {
    "rta_hip_device_count": [_pi],
    "rta_hip_inject_fault": [C.c_char_p, _i],
    "rta_hip_range_push": [C.c_char_p],
    "rta_hip_range_pop": [],
    "rta_hip_set_device": [_i],
    "rta_hip_device_name": [C.c_char_p, _i],
    "rta_hip_malloc": [C.POINTER(_vp), _sz],
    "rta_hip_free": [_vp],
    "rta_hip_memcpy_htod": [_vp, _vp, _sz],
    "rta_hip_memcpy_dtoh": [_vp, _vp, _sz],
    "rta_hip_memcpy_dtod": [_vp, _vp, _sz, _vp],
    "rta_hip_fill_u32": [_vp, _u32, _sz, _vp],
    "rta_hip_stream_create": [C.POINTER(_vp)],
    "rta_hip_stream_destroy": [_vp],
    "rta_hip_stream_synchronize": [_vp],
    "rta_hip_device_synchronize": [],
    "rta_hip_event_create": [C.POINTER(_vp)],
    "rta_hip_event_destroy": [_vp],
    "rta_hip_event_record": [_vp, _vp],
    "rta_hip_stream_wait_event": [_vp, _vp],
    "rta_hip_event_synchronize": [_vp],
    "rta_hip_event_elapsed_ms": [_pf, _vp, _vp],
    "rta_hip_copy_kernel": [_vp, _vp, _sz, _vp],
    "rta_hip_graph_begin": [_vp],
    "rta_hip_graph_end": [_vp, C.POINTER(_vp)],
    "rta_hip_graph_launch": [_vp, _vp],
    "rta_hip_graph_destroy": [_vp],
    "rta_jitter_triangles": [_vp, _vp, _vp, _i, _f, _u64, _u32, _vp, _vp],
    "rta_hip_add_u32": [_vp, _u32, _vp],
    "rta_hip_check_division": [_u64, _u64, C.POINTER(C.c_uint64), _vp],
    "rta_hip_check_division_pairs": [_u32, _u32, C.POINTER(C.c_uint64), _vp],
    "rta_lbvh_init_triangles": [_vp, _vp, _vp, _vp, _vp, _i, _vp],
    "rta_lbvh_voxel_bounds": [_vp, _vp, _i, _vp],
    "rta_lbvh_morton_codes": [_vp, _vp, _vp, _vp, _vp, _i, _vp],
    "rta_lbvh_sort_work_bytes": [_i, _psz],
    "rta_lbvh_sort": [_vp, _vp, _i, _vp, _sz, _vp],
    "rta_lbvh_boxes_work_bytes": [_i, _psz],
    "rta_lbvh_build_tree": [_vp, _vp, _vp, _vp, _i, _vp],
    "rta_lbvh_build_boxes": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "rta_lbvh_permute": [_vp, _vp, _vp, _i, _vp],
    "rta_lbvh_workspace_bytes": [_i, _psz],
    "rta_lbvh_build": [_vp, _i, _vp, _vp, _vp, _sz, _vp, C.POINTER(RtaTimes), _vp],
    "rta_lbvh_build_noalloc": [_vp, _i, _vp, _vp, _vp, _vp],
    "rta_lbvh_build_packed": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp,
                              C.POINTER(RtaTimes), _vp],
    "rta_reset_intersections": [_vp, _i, _i, _vp],
    "rta_setup_shirley_rays": [_vp, _vp, _vp, _f, _f, _i, _i, _vp, _vp, _vp, _f, _vp],
    "rta_trace_cis": [_vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp],
    "rta_bvh_pack": [_vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "rta_trace_cis_packed": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "rta_trace_cis_persistent": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                 _vp, _vp, _vp],
    "rta_trace_cis_persistent_seeded": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp, _vp,
                                        _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "rta_trace_shadow_cis_packed": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                    _vp],
    "rta_trace_shadow_cis_persistent": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp, _vp,
                                        _vp, _vp, _vp, _vp],
    "rta_trace_cis_persistent_ovf": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _vp],
    "rta_trace_shadow_cis_persistent_ovf": [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i, _vp,
                                            _vp, _vp, _vp, _vp, _vp, _vp],
    "rta_bounce_work_bytes": [_i, _psz],
    "rta_diffuse_bounce": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _u64, _vp, _sz, _vp],
    "rta_shadow_rays": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _sz, _vp],
    "rta_batch_create": [_i, _vp, _vp, C.POINTER(_vp)],
    "rta_batch_destroy": [_vp],
    "rta_batch_sizes": [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                        C.POINTER(C.c_longlong)],
    "rta_batch_offsets": [_vp, _vp, _vp, _vp],
    "rta_batch_workspace_bytes": [_vp, _psz],
    "rta_batch_build": [_vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp],
    "rta_batch_pack": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "rta_batch_build_packed": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp],
    "rta_batch_trace_persistent": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp,
                                   _vp],
    "rta_batch_overflow_count": [_vp, _vp, _vp],
    "rta_batch_autotune": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp,
                           _pf, _vp],
    "rta_pred_param_count": [_pi],
    "rta_pred_create": [_i, _u64, C.POINTER(_vp)],
    "rta_pred_destroy": [_vp],
    "rta_pred_get_params": [_vp, _vp],
    "rta_pred_set_params": [_vp, _vp],
    "rta_pred_get_grads": [_vp, _vp],
    "rta_pred_zero_grads": [_vp, _vp],
    "rta_pred_forward": [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "rta_pred_infer": [_vp, _vp, _vp, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    "rta_pred_backward": [_vp, _vp, _vp, _i, _vp, _vp],
    "rta_pred_adam": [_vp, _f, _i, _vp],
    "rta_pred_adam_device": [_vp, _f, _vp, _vp],
    "rta_pred_sample": [_vp, _vp, _vp, _vp, _i, _vp, _u64, _vp, _vp, _vp, _vp],
    "rta_pred_train_step": [_vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _f, _vp, _vp, _vp],
    "rta_pred_train_batches": [_vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _vp, _vp, _f, _vp, _vp,
                               _vp],
    "rta_lbvh_copy_ranges": [_vp, _vp, _i, _vp],
    "rta_treelet_cut": [_vp, _u32, _u32, _vp, _i, _i, _vp, _vp, _vp],
    "rta_treelet_labels": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "rta_trace_seed_treelets": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp,
                                _vp],
    "rta_seed_max_t": [_vp, _vp, _vp, _i, _vp],
}

HOST_SIGNATURES = {
    "rta_scene_icosphere": [_i, _f, _vp, _f, _u64, _vp],
    "rta_scene_heightfield": [_i, _i, _f, _f, _f, _f, _f, _u64, _i, _f, _f, _f, _f, _vp],
    "rta_scene_random_triangles": [_i, _u64, _f, _f, _f, _f, _vp],
    "rta_scene_instances": [_i, _i, _u64, _pi, _vp, _vp],
    "rta_host_shirley_rays": [_vp, _vp, _vp, _f, _f, _i, _i, _vp, _vp, _vp, _f],
    "rta_host_last_error": [],
    "rta_host_scene_bounds": [_vp, _i, _vp, _vp],
    "rta_host_series_camera": [_vp, _vp, _vp, _f, _vp, _vp, _vp],
    "rta_host_axial_positions": [_vp, _vp, _i, _vp],
    "rta_host_ply_vertices": [C.c_char_p, _vp, _i],
    "rta_host_load_obj": [C.c_char_p, _vp, _i],
    "rta_host_median_node_bound": [_i],
    "rta_host_median_build": [_vp, _i, _vp, _vp, _vp],
}


def declared_functions(header: str) -> list[str]:
    """Names of the functions a public header declares (`int rta_...(`)."""
    text = open(os.path.join(INCLUDE_DIR, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rta_\w+)\s*\(", text, flags=re.M)))


def _load(path: str, sigs: dict) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: the HIP extension is not built "
            f"(run `python -c 'import __graft_entry__ as g; g.build()'`)")
    lib = C.CDLL(path)
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    for name in ("rta_hip_last_error", "rta_host_last_error"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = []
            getattr(lib, name).restype = C.c_char_p
    return lib


_hip = None
_host = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        _hip = _load(HIP_LIB_PATH, HIP_SIGNATURES)
    return _hip


def host() -> C.CDLL:
    global _host
    if _host is None:
        _host = _load(HOST_LIB_PATH, HOST_SIGNATURES)
    return _host


def check(status: int, func: str) -> None:
    if status != RTA_OK:
        msg = hip().rta_hip_last_error().decode(errors="replace")
        raise RtaError(status, func, msg)


def call(name: str, *args) -> None:
    """Call a C-ABI entry point of librta_hip and raise RtaError on failure."""
    check(getattr(hip(), name)(*args), name)


def ptr(x) -> C.c_void_p:
    """Device or host address as a ctypes void pointer."""
    if x is None:
        return C.c_void_p(None)
    if isinstance(x, np.ndarray):
        return C.c_void_p(x.ctypes.data)
    if hasattr(x, "ptr"):
        return C.c_void_p(x.ptr)
    return C.c_void_p(int(x))
