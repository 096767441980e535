"""Batches of meshes on the GPU: every mesh of a batch built in shared
launches and all rays traced by one persistent launch must equal the
oracle's build and traversal of that mesh alone."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd import scenes
from rta_amd.abi import TRI_DTYPE
from rta_amd.batch import MeshBatch

F32MAX = np.float32(np.finfo(np.float32).max)


def say(msg):
    print(f"[batch] {msg}", flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _clustered(n, cells, seed):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, TRI_DTYPE)
    centres = rng.uniform(0, 1, size=(cells, 3)).astype(np.float32)
    base = centres[rng.integers(0, cells, size=n)]
    for k in ("a", "b", "c"):
        t[k] = base + rng.uniform(-1e-5, 1e-5, size=(n, 3)).astype(np.float32)
    t[0]["a"], t[0]["b"], t[0]["c"] = (0, 0, 0), (1e-4, 0, 0), (0, 1e-4, 1e-4)
    t[1]["a"], t[1]["b"], t[1]["c"] = (1, 1, 1), (0.99, 1, 1), (1, 0.99, 0.99)
    return t


def _edge_batch():
    """Mesh sizes at the stage boundaries: 1 and 2 triangles, one below, at
    and above a 1024-leaf chunk, two merge levels (> 65,536 leaves),
    duplicate-code runs, plus a few ordinary meshes; one mesh without rays."""
    rng = np.random.default_rng(5)
    meshes = [scenes.random_triangles(1, seed=1), scenes.random_triangles(2, seed=2),
              scenes.random_triangles(1023, seed=3), scenes.random_triangles(1024, seed=4),
              scenes.random_triangles(1025, seed=5), scenes.heightfield(187, 187, seed=6),
              _clustered(3000, 17, seed=7), scenes.icosphere(3, jitter=0.05, seed=8),
              scenes.heightfield(20, 31, seed=9)]
    meshes += scenes.batch_meshes(12, seed=10, tris=(200, 20_000), rays=(500, 4000))[0]
    rays = [scenes.mesh_rays(m, int(rng.integers(300, 3000)), rng) for m in meshes]
    rays[3] = (np.zeros(0, np.float32),) * 3
    return meshes, rays


def _check_mesh(b, m, mesh, ray_set, shadow_hits=None):
    got = b.mesh_results(m)
    exp = oracle.lbvh_build(mesh)
    assert np.array_equal(got["indices"], exp["indices"]), f"mesh {m} indices"
    assert np.array_equal(bits(got["tris"]), bits(exp["tris"])), f"mesh {m} tris"
    gn, en = got["nodes"], exp["nodes"]
    for f in ("type_left_elems", "right_tris", "min", "max"):
        assert np.array_equal(gn[f], en[f]), f"mesh {m} nodes.{f}"
    o, d, t = ray_set
    if len(t) == 0:
        return
    ref, deep = oracle.trace_kernel(exp["tris"], exp["nodes"], o, d, t)
    assert deep == 0
    g = got["hits"]
    hit = ref["t"] != F32MAX
    assert np.array_equal(g["t"] != F32MAX, hit), f"mesh {m} hit flags"
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(g[f][hit]), bits(ref[f][hit])), f"mesh {m} {f}"
    assert np.array_equal(g["ref"][hit], ref["ref"][hit]), f"mesh {m} ref"


@pytest.mark.gpu
def test_batch_build_and_trace_bit_exact(gpu):
    meshes, rays = _edge_batch()
    say(f"{len(meshes)} meshes, {sum(len(m) for m in meshes)} triangles")
    b = MeshBatch(meshes, rays)
    b.build()
    b.trace()
    b.stream.synchronize()
    b.check_errors()
    assert int(b.status.download(np.uint32, 1)[0]) & ~1 == 0
    for m in range(len(meshes)):
        _check_mesh(b, m, meshes[m], rays[m])


@pytest.mark.gpu
def test_batch_shadow_trace_bit_exact(gpu):
    meshes, rays = _edge_batch()
    b = MeshBatch(meshes, rays)
    b.build()
    b.trace(any_hit=True)
    b.stream.synchronize()
    b.check_errors()
    for m in (0, 2, 5, 6, 7, 10, 15):
        got = b.mesh_results(m)
        exp = oracle.lbvh_build(meshes[m])
        o, d, t = rays[m]
        ref, _ = oracle.trace_kernel(exp["tris"], exp["nodes"], o, d, t, any_hit=True)
        hit = ref["t"] != F32MAX
        g = got["hits"]
        assert np.array_equal(g["t"] != F32MAX, hit)
        assert np.array_equal(g["ref"][hit], ref["ref"][hit])
        assert np.array_equal(bits(g["t"][hit]), bits(ref["t"][hit]))


@pytest.mark.gpu
def test_batch_rebuild_and_autotune(gpu):
    """A second build over the same buffers gives the same trees, and the
    autotuned launch shape the same hits."""
    meshes, rays = _edge_batch()
    b = MeshBatch(meshes, rays)
    b.build()
    b.trace()
    b.stream.synchronize()
    first = [b.mesh_results(m) for m in range(len(meshes))]
    b.build()
    cfg, ms = b.autotune(rounds=1, reps=1)
    say(f"autotuned {cfg.as_tuple()} {ms:.3f} ms")
    assert cfg.block in (64, 128) and 1 <= cfg.leaf_min <= 64 and ms > 0
    b.trace(config=cfg)
    b.stream.synchronize()
    b.check_errors()
    for m in range(len(meshes)):
        g = b.mesh_results(m)
        assert np.array_equal(bits(g["nodes"]), bits(first[m]["nodes"]))
        if g["hits"] is not None:
            assert np.array_equal(bits(g["hits"]), bits(first[m]["hits"]))


def test_batch_rejects_bad_sizes():
    import ctypes as C
    from rta_amd.abi import RtaError, call, ptr
    h = C.c_void_p()
    with pytest.raises(RtaError):
        call("rta_batch_create", 0, None, None, C.byref(h))
    sizes = np.array([10, 0, 5], np.int32)
    with pytest.raises(RtaError):
        call("rta_batch_create", 3, ptr(sizes), None, C.byref(h))
    sizes = np.full(257, 10, np.int32)
    with pytest.raises(RtaError):
        call("rta_batch_create", 257, ptr(sizes), None, C.byref(h))
