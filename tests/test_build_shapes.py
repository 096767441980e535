"""LBVH builds of scenes that stress the bottom-up pass, which also emits the
hierarchy: every triangle identical (the whole tree collapses into one
leaf at the root), a few distinct triangles repeated (thousands of merged
leaves), positions crowded toward one corner at geometric spacing (long
common prefixes, deep spines across chunk boundaries), and sizes at the
chunk and merge-group boundaries (1,024 leaves per chunk, 16 chunks per
merge block, so 16,384 and 262,144 leaves fill one and two merge levels).  Both full builds (rta_lbvh_build and rta_lbvh_build_packed)
against the oracle (lbvh-kernels.cu:186-341 restated): nodes, sorted
permutation, reordered triangles and status; the packed layout against
rta_bvh_pack of the plain build."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from rta_amd import scenes
from rta_amd.abi import TRI_DTYPE

from test_gpu_parity import _build_packed, _build_then_pack


def say(msg):
    print(f"[build-shapes] {msg}", flush=True)
    sys.stdout.flush()


def identical(n):
    return np.repeat(scenes.random_triangles(1, seed=3), n)


def repeated(n, distinct, seed):
    base = scenes.random_triangles(distinct, seed=seed)
    return base[np.random.default_rng(seed).integers(0, distinct, n)]


def corner_crowd(n, seed):
    """Small triangles at 0.5^k (k = 0 .. 23 per axis, drawn independently)
    plus jitter of 1e-9: most share long Morton prefixes with their
    neighbours."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 24, size=(n, 3))
    c = (0.5 ** k + rng.uniform(0, 1e-9, size=(n, 3))).astype(np.float32)
    t = np.zeros(n, TRI_DTYPE)
    e = np.float32(1e-4)
    t["a"] = c
    t["b"] = c + np.array([e, 0, 0], np.float32)
    t["c"] = c + np.array([0, e, 0], np.float32)
    return t


SHAPES = {
    "identical_3000": lambda: identical(3000),
    "repeated_50_of_20000": lambda: repeated(20_000, 50, 4),
    "repeated_5000_of_70000": lambda: repeated(70_000, 5000, 5),
    "corner_crowd_60000": lambda: corner_crowd(60_000, 6),
    "chunk_8192": lambda: scenes.random_triangles(8192, seed=7),
    "chunk_8193": lambda: scenes.random_triangles(8193, seed=8),
    "merge_65536": lambda: scenes.random_triangles(65_536, seed=9),
    "merge_65537": lambda: scenes.random_triangles(65_537, seed=10),
    "merge_16384": lambda: scenes.random_triangles(16_384, seed=12),
    "merge_16385": lambda: scenes.random_triangles(16_385, seed=13),
    "merge_262144": lambda: scenes.random_triangles(262_144, seed=14),
    "merge_262145": lambda: scenes.random_triangles(262_145, seed=15),
}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SHAPES))
def test_build_shapes_against_the_oracle(gpu, name):
    tris = SHAPES[name]()
    n = len(tris)
    ref = oracle.lbvh_build(tris)
    a, b = _build_then_pack(tris), _build_packed(tris)
    inner = (ref["nodes"]["type_left_elems"] & 1) == 1
    say(f"{name}: n={n}, {n - 1 - int(inner[:n - 1].sum())} merged inner nodes, "
        f"status {ref['status']}")
    for got in (a, b):
        assert got["nodes"].tobytes() == ref["nodes"].tobytes(), name
        assert np.array_equal(got["idx"], ref["indices"]), name
        assert np.array_equal(bits(got["tris"]), bits(ref["tris"])), name
        assert got["status"] == ref["status"], name
    assert np.array_equal(a["tri4"], b["tri4"]) and a["flags"] == b["flags"]
    assert np.array_equal(a["pairs"][inner], b["pairs"][inner])
    if name.startswith("identical"):
        root = ref["nodes"][0]
        assert root["type_left_elems"] == 2 * n and root["right_tris"] == 0


def test_shapes_on_the_oracle():
    """The shapes do what they are for, on the CPU: the identical scene's
    root is one leaf of every triangle, the repeated scenes merge thousands
    of inner nodes, and the corner crowd builds a tree deeper than 24 (a
    balanced tree of its 60,000 leaves is 16 deep)."""
    ref = oracle.lbvh_build(identical(3000))
    assert ref["nodes"][0]["type_left_elems"] == 6000
    ref = oracle.lbvh_build(repeated(20_000, 50, 4))
    merged = int(((ref["nodes"]["type_left_elems"][:19_999] & 1) == 0).sum())
    assert merged > 19_000
    t = corner_crowd(60_000, 6)
    ref = oracle.lbvh_build(t)
    nodes = ref["nodes"]
    n = len(t)
    depth = np.zeros(2 * n - 1, np.int32)
    stack = [0]
    deepest = 0
    while stack:
        i = stack.pop()
        w = int(nodes[i]["type_left_elems"])
        if w & 1 and i < n - 1:
            for c in (w >> 1, int(nodes[i]["right_tris"])):
                depth[c] = depth[i] + 1
                deepest = max(deepest, depth[c])
                stack.append(c)
    assert deepest > 24, deepest
