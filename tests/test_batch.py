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
from rta_amd.abi import HIT_DTYPE, TRI_DTYPE
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
def test_batch_build_packed_equals_build_then_pack(gpu):
    """rta_batch_build_packed writes the traversal layout in the build's own
    passes: nodes, indices, reordered triangles and status equal
    rta_batch_build's; tri4, the per-mesh flags and every inner node's pair
    record equal rta_batch_pack's; and the hits (which also read the
    per-mesh root words) are identical."""
    from rta_amd.abi import NODE_DTYPE
    meshes, rays = _edge_batch()
    out = []
    for packed in (False, True):
        b = MeshBatch(meshes, rays)
        b.tri4.fill_u32(0x7fc00000)           # every tri4 word must be written
        b.flags.fill_u32(0xffffffff)          # and every flag cleared
        b.stream.synchronize()
        b.build(packed=packed)
        b.trace()
        b.stream.synchronize()
        b.check_errors()
        out.append({
            "nodes": b.nodes.download(NODE_DTYPE, b.total_nodes),
            "indices": b.indices.download(np.uint32, b.total_tris),
            "tris": b.tris.download(np.uint8, 104 * b.total_tris),
            "status": int(b.status.download(np.uint32, 1)[0]),
            "tri4": b.tri4.download(np.uint8, 48 * b.total_tris),
            "flags": b.flags.download(np.uint32, b.M),
            "pairs": b.pairs.download(np.uint8, 64 * b.total_nodes).reshape(-1, 64),
            "hits": b.hits.download(np.uint8, 16 * b.total_rays),
        })
    a, c = out
    assert a["nodes"].tobytes() == c["nodes"].tobytes()
    for k in ("indices", "tris", "tri4", "flags", "hits"):
        assert np.array_equal(a[k], c[k]), k
    assert a["status"] == c["status"]
    inner = (a["nodes"]["type_left_elems"] & 1) == 1
    assert np.array_equal(a["pairs"][inner], c["pairs"][inner])
    say(f"packed batch build equals build + pack ({int(inner.sum())} pair records)")


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


@pytest.mark.gpu
@pytest.mark.parametrize("stack", [16, 24])
def test_short_stack_traversal_equals_full_stack(gpu, stack):
    """Batch traversal with a 16- or 24-entry LDS stack: rays that would
    push beyond it are traced again with 32 entries, so every record equals
    the 32-entry traversal's bit for bit.  Two hand-made chain meshes (the
    trees of test_stack_depth: 20 and 28 pushes) force such rays; the
    overflow count equals the rays whose oracle traversal holds more than
    `stack` entries, and the chain records equal the oracle's."""
    import ctypes as C
    from rta_amd.abi import TraceConfig, call, ptr
    from test_stack_depth import chain_rays, chain_tree

    meshes, rays = _edge_batch()
    chains = [chain_tree(20), chain_tree(28)]
    for j, (tris_c, _) in enumerate(chains):
        meshes.append(tris_c)
        rays.append(chain_rays(n_random=97 + j, seed=11 + j))
    b = MeshBatch(meshes, rays)
    b.build()
    b.stream.synchronize()              # the uploads below run on another stream
    s = ptr(b.stream)
    # the chain meshes get their hand-made trees instead of LBVH ones
    for j, (tris_c, nodes_c) in enumerate(chains):
        m = len(meshes) - 2 + j
        call("rta_hip_memcpy_htod", C.c_void_p(b.tris.ptr + 104 * int(b.tri_off[m])),
             ptr(tris_c), C.c_size_t(tris_c.nbytes))
        call("rta_hip_memcpy_htod", C.c_void_p(b.nodes.ptr + 32 * int(b.node_off[m])),
             ptr(nodes_c), C.c_size_t(nodes_c.nbytes))
    call("rta_batch_pack", b.handle, ptr(b.pairs), ptr(b.tri4), ptr(b.flags), ptr(b.nodes),
         ptr(b.tris), s)
    b.trace(config=TraceConfig(64, 16, 32, 0, 10, 32))
    b.stream.synchronize()
    b.check_errors()
    assert b.overflow_count() == 0
    full = b.hits.download(np.uint8, 16 * b.total_rays).copy()
    b.trace(config=TraceConfig(64, 16, 32, 0, 10, stack))
    b.stream.synchronize()
    b.check_errors()
    n_ovf = b.overflow_count()
    assert np.array_equal(b.hits.download(np.uint8, 16 * b.total_rays), full)
    expected = 0
    for m in range(len(meshes)):
        o, d, t = rays[m]
        if len(t) == 0:
            continue
        if m >= len(meshes) - 2:
            tris_c, nodes_c = chains[m - (len(meshes) - 2)]
            ref, deep, depth = oracle.trace_kernel(tris_c, nodes_c, o, d, t, depth=True)
            g = b.mesh_results(m)["hits"]
            for f in ("t", "beta", "gamma", "ref"):
                assert np.array_equal(bits(g[f]), bits(ref[f])), (m, f)
        else:
            exp = oracle.lbvh_build(meshes[m])
            _, deep, depth = oracle.trace_kernel(exp["tris"], exp["nodes"], o, d, t, depth=True)
        assert deep == 0
        expected += int(np.count_nonzero(depth > stack))
    say(f"stack {stack}: {n_ovf} of {b.total_rays} rays traced again")
    assert n_ovf == expected and n_ovf >= (len(rays[-1][2]) if stack == 24 else
                                            len(rays[-1][2]) + len(rays[-2][2]))


@pytest.mark.gpu
def test_batch_frame_graph_replays_equal_eager(gpu):
    """Build + pack + trace of the whole batch captured as one graph (with a
    short stack, so the overflow pass is captured too): every replay, with
    eager launches and synchronisations between replays, gives the eager
    frame's trees and hits bit for bit."""
    from rta_amd.abi import TraceConfig
    meshes, rays = _edge_batch()
    b = MeshBatch(meshes, rays)
    cfg = TraceConfig(128, 16, 32, 0, 8, 16)
    b.build()
    b.trace(config=cfg)
    b.stream.synchronize()
    b.check_errors()
    eager_nodes = b.nodes.download(np.uint8, 32 * b.total_nodes).copy()
    eager = b.hits.download(HIT_DTYPE, b.total_rays).copy()
    hit = eager["t"] != F32MAX
    g = b.capture(config=cfg)
    for rep in range(3):
        # eager work between replays: scribble over the outputs
        b.hits.fill_u32(0x7f7f7f7f, stream=b.stream)
        b.nodes.fill_u32(0, stream=b.stream)
        b.stream.synchronize()
        g.launch(b.stream)
        b.stream.synchronize()
        b.check_errors()
        assert np.array_equal(b.nodes.download(np.uint8, 32 * b.total_nodes), eager_nodes), rep
        got = b.hits.download(HIT_DTYPE, b.total_rays)
        # the reset defines t, beta and gamma of a miss; ref only for a hit
        for f in ("t", "beta", "gamma"):
            assert np.array_equal(bits(got[f]), bits(eager[f])), (rep, f)
        assert np.array_equal(got["ref"][hit], eager["ref"][hit]), rep
    say(f"3 graph replays equal the eager frame ({int(hit.sum())} hits)")


def test_short_stack_rejected_outside_batches():
    """rta_trace_cis_persistent holds no overflow list (its _ovf form does):
    stack 16 or 24 is an argument error there, and 20 anywhere (checked
    before any device work)."""
    import ctypes as C
    from rta_amd.abi import RTA_E_INVALID, TraceConfig
    lib = rta_amd.hip()
    fake = C.c_void_p(64)
    for stack in (16, 24, 20):
        cfg = TraceConfig(64, 16, 32, 0, 10, stack)
        rc = lib.rta_trace_cis_persistent(fake, fake, fake, 1, 0, fake, fake, fake, 10, None,
                                          fake, fake, fake, C.byref(cfg), None)
        assert rc == RTA_E_INVALID
        assert b"stack" in lib.rta_hip_last_error()


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
