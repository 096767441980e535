"""Scenes at the edges of the reference's arithmetic.

Planar scenes: one axis of the scene box has zero extent, so the Morton
scale on that axis is 1024/0 = inf and every centroid offset on it is
0 * inf = NaN, which the conversion turns into 0 (lbvh-kernels.cu:177-180,
415-417): all codes have that axis's bits clear.

NaN vertices: a triangle's box takes a NaN coordinate only where the NaN
comes last in MIN(MIN(a, b), c) (lbvh-kernels.cu:141-153); its centroid is
NaN and its code bits on that axis 0.  The scene box skips NaN coordinates
(the reference's thrust::reduce over MIN/MAX is order-dependent once a NaN
is an operand, so it has no single answer there); boxes of inner nodes
follow MIN(left, right), NaN included; triangles with a NaN vertex are never
hit.  Tracing such a scene takes the IEEE slab test (NaN coordinates are
outside the exact-division range).

CPU tests pin the oracle on these scenes; GPU tests compare every stage and
the traversal with it."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd import scenes
from rta_amd.abi import NODE_DTYPE, TRI_DTYPE
from rta_amd.device import DeviceBuffer
from helpers import radix_tree

F32MAX = np.float32(np.finfo(np.float32).max)
NAN = np.float32(np.nan)
Z_BITS = sum(1 << (3 * i + 2) for i in range(10))
X_BITS = sum(1 << (3 * i) for i in range(10))


def say(msg):
    print(f"[edge-scenes] {msg}", flush=True)
    sys.stdout.flush()


def planar_z(n_random=3000, grid=40, seed=3):
    g = scenes.heightfield(grid, grid, z0=0.3, amplitude=0.0, seed=1)
    r = scenes.random_triangles(n_random, seed=seed)
    for k in ("a", "b", "c"):
        r[k][:, 2] = np.float32(0.3)
    return np.concatenate([g, r])


def planar_x():
    t = planar_z(seed=4)
    for k in ("a", "b", "c"):
        t[k] = t[k][:, ::-1].copy()
    return t


def nan_scene():
    t = np.concatenate([scenes.heightfield(40, 40, seed=5), scenes.random_triangles(2000, seed=6)])
    t[5]["a"][0] = NAN            # dropped by MIN(MIN(a, b), c): finite box
    t[17]["b"][1] = NAN
    t[33]["c"][2] = NAN           # last operand: the box's z bounds are NaN
    t[60]["a"] = t[60]["b"] = t[60]["c"] = (NAN, NAN, NAN)
    t[99]["c"] = (NAN, NAN, NAN)
    t[len(t) - 1]["c"][0] = NAN   # the last triangle of the scene
    return t


SCENES = {"planar_z": planar_z, "planar_x": planar_x, "nan_vertices": nan_scene}


def sel_min(a, b):
    return np.where(a < b, a, b)


def sel_max(a, b):
    return np.where(a > b, a, b)


def same(a, b):
    """float arrays equal by value, NaN equal to NaN"""
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def rays(n, seed, lo=-0.2, hi=1.2):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[: n // 10, rng.integers(0, 3)] = 0.0          # some axis-parallel rays
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o.reshape(-1), d.reshape(-1), np.full(n, F32MAX, np.float32)


# ------------------------------------------------------------- CPU pins

@pytest.mark.parametrize("name,bits", [("planar_z", Z_BITS), ("planar_x", X_BITS)])
def test_planar_codes_and_topology_on_oracle(name, bits):
    tris = SCENES[name]()
    ref = oracle.lbvh_build(tris)
    axis = 2 if name == "planar_z" else 0
    b = ref["bounds"][0]
    assert b["min"][axis] == b["max"][axis]
    # 1024/0 = inf, 0 * inf = NaN -> 0: the flat axis contributes no bits
    assert np.all(ref["codes_input"] & np.uint32(bits) == 0)
    assert np.any(ref["codes_input"] != 0)
    # the emitted hierarchy is the binary radix tree over (code, position)
    n = len(tris)
    links = ref["nodes"][: n - 1]
    tree = radix_tree(ref["codes_sorted"])
    dup = np.diff(ref["codes_sorted"].astype(np.int64)) == 0
    if not dup.any():
        for i, (_, _, left, right) in tree.items():
            assert links[i]["type_left_elems"] >> 1 == left and links[i]["right_tris"] == right


def test_planar_traversal_on_oracle_against_brute_force():
    tris = planar_z()
    ref = oracle.lbvh_build(tris)
    o, d, t = rays(4000, seed=7)
    hk, deep = oracle.trace_kernel(ref["tris"], ref["nodes"], o, d, t)
    hb = oracle.trace_brute(ref["tris"], o, d)
    assert deep == 0
    hit_b = hb["t"] != F32MAX
    hit_k = hk["t"] != F32MAX
    assert hit_b.sum() > 500
    # the BVH's t is never below the brute-force minimum and almost always
    # equal to it; it can be a few ulps above where the ray crosses two
    # coplanar triangles and the computed entry distance of the nearer one's
    # zero-thickness box exceeds the t already found (the kernel prunes
    # boxes with dist >= closest t); it misses a brute-force hit only where
    # a ray grazes a box boundary
    both = hit_b & hit_k
    tk, tb = hk["t"][both].astype(np.float64), hb["t"][both].astype(np.float64)
    assert np.all(tk >= tb) and np.all((tk - tb) / tb < 2.0 ** -16)
    assert (tk == tb).mean() > 0.99
    assert (hit_b & ~hit_k).sum() <= 0.01 * hit_b.sum()
    assert not (hit_k & ~hit_b).any()


def test_nan_scene_on_oracle():
    tris = nan_scene()
    ref = oracle.lbvh_build(tris)
    v = np.stack([tris["a"], tris["b"], tris["c"]], axis=1)
    lo = sel_min(sel_min(v[:, 0], v[:, 1]), v[:, 2])
    hi = sel_max(sel_max(v[:, 0], v[:, 1]), v[:, 2])
    # the scene box: min/max over the non-NaN box coordinates
    assert np.array_equal(ref["bounds"][0]["min"], np.nanmin(lo, axis=0))
    assert np.array_equal(ref["bounds"][0]["max"], np.nanmax(hi, axis=0))
    assert np.isnan(lo[33, 2]) and not np.isnan(lo[5, 0])
    ctr = ((tris["a"] + tris["b"]) + tris["c"]) / np.float32(3)
    nanc = np.isnan(ctr)
    for axis, bits in ((0, X_BITS), (1, X_BITS << 1), (2, Z_BITS)):
        assert np.all(ref["codes_input"][nanc[:, axis]] & np.uint32(bits) == 0)
    # every triangle appears once; no ray hits a triangle with a NaN vertex
    assert np.array_equal(np.sort(ref["indices"]), np.arange(len(tris), dtype=np.uint32))
    o, d, t = rays(4000, seed=8)
    h, deep = oracle.trace_kernel(ref["tris"], ref["nodes"], o, d, t)
    assert deep == 0
    hit = h["t"] != F32MAX
    assert hit.sum() > 500
    bad = np.isnan(np.stack([ref["tris"]["a"], ref["tris"]["b"], ref["tris"]["c"]],
                            axis=1)).any(axis=(1, 2))
    assert not bad[h["ref"][hit]].any()


# ------------------------------------------------------------- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SCENES))
def test_edge_scene_stages_bit_exact(gpu, name):
    from test_gpu_parity import bits, gpu_stages
    tris = SCENES[name]()
    say(f"stages scene={name} n={len(tris)}")
    ref = oracle.lbvh_build(tris)
    got = gpu_stages(tris)
    assert np.array_equal(bits(got["scratch"]), bits(tris))
    v = np.stack([tris["a"], tris["b"], tris["c"]], axis=1)
    assert same(got["boxes"]["min"], sel_min(sel_min(v[:, 0], v[:, 1]), v[:, 2]))
    assert same(got["boxes"]["max"], sel_max(sel_max(v[:, 0], v[:, 1]), v[:, 2]))
    assert np.array_equal(got["bounds"]["min"], ref["bounds"]["min"])
    assert np.array_equal(got["bounds"]["max"], ref["bounds"]["max"])
    assert np.array_equal(got["codes_input"], ref["codes_input"])
    assert got["status"] == ref["status"]
    assert np.array_equal(got["indices"], ref["indices"])
    assert np.array_equal(got["parents"][1:], ref["parents"][1:])
    for f in ("type_left_elems", "right_tris"):
        assert np.array_equal(got["nodes"][f], ref["nodes"][f])
    assert same(got["nodes"]["min"], ref["nodes"]["min"])
    assert same(got["nodes"]["max"], ref["nodes"]["max"])
    assert np.array_equal(bits(got["tris"]), bits(ref["tris"]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("layout", ["plain", "packed", "persistent"])
def test_edge_scene_trace_bit_exact(gpu, name, layout):
    from test_gpu_parity import assert_hits_equal, trace_gpu
    tris = SCENES[name]()
    say(f"trace scene={name} layout={layout}")
    ref_b = oracle.lbvh_build(tris)
    o, d, t = rays(20_000, seed=9)
    ref, deep = oracle.trace_kernel(ref_b["tris"], ref_b["nodes"], o, d, t)
    assert deep == 0
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    if layout != "plain":
        # zero-thickness boxes stay on the exact-division path; NaN
        # coordinates send the whole scene to the IEEE one
        assert bvh.packed().exact_division_path() == (name != "nan_vertices")
    got, _ = trace_gpu(bvh, o, d, t, layout)
    assert_hits_equal(got, ref)
