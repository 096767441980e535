"""ORACLE -- test infrastructure only (the checker, never the product).

ctypes wrapper of oracle/_build/liboracle.so, the CPU restatement of the
reference's BVH path (see oracle/oracle.cpp for the cited reference lines).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

# dtypes mirror include/rta_hip.h (kept local so the oracle does not import
# the product package)
TRI_DTYPE = np.dtype(
    [("a", "<f4", 3), ("b", "<f4", 3), ("c", "<f4", 3),
     ("na", "<f4", 3), ("nb", "<f4", 3), ("nc", "<f4", 3),
     ("ta", "<f4", 2), ("tb", "<f4", 2), ("tc", "<f4", 2),
     ("material_index", "<i4"), ("_pad", "<i4")])
AABB_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3)])
NODE_DTYPE = np.dtype([("type_left_elems", "<u4"), ("right_tris", "<u4"),
                       ("min", "<f4", 3), ("max", "<f4", 3)])
HIT_DTYPE = np.dtype([("t", "<f4"), ("beta", "<f4"), ("gamma", "<f4"), ("ref", "<u4")])
F32MAX = np.float32(np.finfo(np.float32).max)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i, u32, u64, d = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_double
        sigs = {
            "oracle_intersect_tri": (i, [vp, vp, vp, vp]),
            "oracle_intersect_aabb": (i, [vp, vp, vp, vp]),
            "oracle_morton_encode": (u32, [u32, u32, u32]),
            "oracle_lbvh_build": (i, [vp, i, vp, vp, vp, vp, vp, vp, vp]),
            "oracle_median_build": (i, [vp, i, vp, vp, vp]),
            "oracle_trace_cis": (d, [vp, vp, vp, vp, vp, i, i, vp, vp, vp, i]),
            "oracle_trace_dis": (d, [vp, vp, vp, vp, vp, i, i, vp, i]),
            "oracle_trace_kernel": (i, [vp, vp, vp, vp, vp, i, vp, i, vp, vp, vp, i]),
            "oracle_trace_brute": (None, [vp, i, vp, vp, i, vp]),
            "oracle_diffuse_bounce": (i, [vp, vp, vp, vp, i, u64, vp, vp, vp, vp]),
            "oracle_shadow_rays": (i, [vp, vp, vp, vp, i, vp, vp, vp, vp, vp]),
            "oracle_jitter_triangles": (None, [vp, vp, vp, i, C.c_float, u64, u32]),
            "oracle_max_threads": (i, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def intersect_tri(tri, o, d):
    """(hit, t, beta, gamma) of intersect_tri_opt for one triangle/ray."""
    t = np.ascontiguousarray(tri).reshape(1)
    o, d = _f32(o), _f32(d)
    out = np.zeros(3, np.float32)
    hit = lib().oracle_intersect_tri(_p(t), _p(o), _p(d), _p(out))
    return bool(hit), out[0], out[1], out[2]


def intersect_aabb(box_min, box_max, o, d):
    """(hit, dist) of intersect_aabb."""
    b = np.zeros(1, AABB_DTYPE)
    b["min"], b["max"] = box_min, box_max
    o, d = _f32(o), _f32(d)
    dist = np.zeros(1, np.float32)
    hit = lib().oracle_intersect_aabb(_p(b), _p(o), _p(d), _p(dist))
    return bool(hit), dist[0]


def morton_encode(x: int, y: int, z: int) -> int:
    return int(lib().oracle_morton_encode(x, y, z))


def lbvh_build(tris: np.ndarray) -> dict:
    n = len(tris)
    tris = np.ascontiguousarray(tris)
    out = {
        "tris": np.zeros(n, TRI_DTYPE), "nodes": np.zeros(2 * n - 1, NODE_DTYPE),
        "indices": np.zeros(n, np.uint32), "codes_sorted": np.zeros(n, np.uint32),
        "codes_input": np.zeros(n, np.uint32), "parents": np.zeros(2 * n - 1, np.uint32),
        "bounds": np.zeros(1, AABB_DTYPE),
    }
    st = lib().oracle_lbvh_build(_p(tris), n, _p(out["tris"]), _p(out["nodes"]),
                                 _p(out["indices"]), _p(out["codes_sorted"]),
                                 _p(out["codes_input"]), _p(out["parents"]), _p(out["bounds"]))
    if st < 0:
        raise ValueError("oracle_lbvh_build: n must be > 0")
    out["status"] = st
    return out


def median_build(tris: np.ndarray) -> dict:
    n = len(tris)
    tris = np.ascontiguousarray(tris)
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    tout = np.zeros(n, TRI_DTYPE)
    order = np.zeros(n, np.uint32)
    nn = lib().oracle_median_build(_p(tris), n, _p(nodes), _p(tout), _p(order))
    return {"nodes": nodes[:nn].copy(), "tris": tout, "order": order}


def trace_cis(tris, nodes, o, d, t, w, h=1, stats=False, threads=0, first=False):
    """Child-is closest hit (bbvh/bbvh.h:407-455).  Returns (hits, ms, stats),
    plus the first accepted record per ray when first=True."""
    n = w * h
    hits = np.zeros(n, HIT_DTYPE)
    st = np.zeros(2 * n, np.uint32) if stats else None
    fr = np.zeros(n, HIT_DTYPE) if first else None
    tris, nodes = np.ascontiguousarray(tris), np.ascontiguousarray(nodes)
    o, d, t = _f32(o), _f32(d), _f32(t)
    ms = lib().oracle_trace_cis(_p(tris), _p(nodes), _p(o), _p(d), _p(t), w, h, _p(hits), _p(st),
                                _p(fr), threads)
    return (hits, ms, st, fr) if first else (hits, ms, st)


def trace_kernel(tris, nodes, o, d, t, any_hit=False, hits_in=None, threads=0, bound=None,
                 stats=False, depth=False):
    """The GPU tracers' semantics (kernels.cu:777-855, shadow :857-939).
    hits_in: the records the rays start from (default: the reset sentinel).
    bound: optional per-ray inclusive bound on child-box entry distances.
    Returns (hits, rays that needed more than 32 stack entries), then with
    stats=True per-ray (inner-node visits, triangle tests), then with
    depth=True each ray's largest number of simultaneous stack entries."""
    n = len(t)
    if hits_in is None:
        hits = np.zeros(n, HIT_DTYPE)
        hits["t"], hits["beta"], hits["gamma"] = F32MAX, 1.0, 2.0
    else:
        hits = np.array(hits_in, dtype=HIT_DTYPE, copy=True)
    tris, nodes = np.ascontiguousarray(tris), np.ascontiguousarray(nodes)
    o, d, t = _f32(o), _f32(d), _f32(t)
    b = None if bound is None else _f32(bound)
    st = np.zeros(2 * n, np.uint32) if stats else None
    dp = np.zeros(n, np.uint32) if depth else None
    deep = lib().oracle_trace_kernel(_p(tris), _p(nodes), _p(o), _p(d), _p(t), n, _p(hits),
                                     int(bool(any_hit)), _p(b), _p(st), _p(dp), threads)
    out = (hits, int(deep))
    if stats:
        out += (st.reshape(-1, 2),)
    if depth:
        out += (dp,)
    return out


def trace_dis(tris, nodes, o, d, t, w, h=1, threads=0):
    n = w * h
    hits = np.zeros(n, HIT_DTYPE)
    tris, nodes = np.ascontiguousarray(tris), np.ascontiguousarray(nodes)
    o, d, t = _f32(o), _f32(d), _f32(t)
    ms = lib().oracle_trace_dis(_p(tris), _p(nodes), _p(o), _p(d), _p(t), w, h, _p(hits), threads)
    return hits, ms


def trace_brute(tris, o, d):
    n = len(o) // 3
    hits = np.zeros(n, HIT_DTYPE)
    tris = np.ascontiguousarray(tris)
    o, d = _f32(o), _f32(d)
    lib().oracle_trace_brute(_p(tris), len(tris), _p(o), _p(d), n, _p(hits))
    return hits


def diffuse_bounce(hits, tris, o, d, seed):
    n = len(hits)
    oo = np.zeros(3 * n, np.float32)
    od = np.zeros(3 * n, np.float32)
    ot = np.zeros(n, np.float32)
    src = np.zeros(n, np.uint32)
    hits, tris = np.ascontiguousarray(hits), np.ascontiguousarray(tris)
    o, d = _f32(o), _f32(d)
    k = lib().oracle_diffuse_bounce(_p(hits), _p(tris), _p(o), _p(d), n, C.c_uint64(seed), _p(oo),
                                    _p(od), _p(ot), _p(src))
    return oo[:3 * k], od[:3 * k], ot[:k], src[:k]


def shadow_rays(hits, tris, o, d, light):
    n = len(hits)
    oo = np.zeros(3 * n, np.float32)
    od = np.zeros(3 * n, np.float32)
    ot = np.zeros(n, np.float32)
    src = np.zeros(n, np.uint32)
    hits, tris = np.ascontiguousarray(hits), np.ascontiguousarray(tris)
    o, d, lp = _f32(o), _f32(d), _f32(light)
    k = lib().oracle_shadow_rays(_p(hits), _p(tris), _p(o), _p(d), n, _p(lp), _p(oo), _p(od),
                                 _p(ot), _p(src))
    return oo[:3 * k], od[:3 * k], ot[:k], src[:k]


def jitter_triangles(verts, idx, sigma, seed, frame):
    n = len(idx) // 3
    out = np.zeros(n, TRI_DTYPE)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    lib().oracle_jitter_triangles(_p(out), _p(verts), _p(idx), n, float(sigma),
                                  C.c_uint64(seed), frame)
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())
