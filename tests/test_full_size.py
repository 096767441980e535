"""The benchmark workload at its full size against the oracle: the
2,297,152-triangle heightfield + random scene, its LBVH, 1920x1080 primary
rays, the diffuse bounce rays generated on the GPU and their closest hits,
every record bit for bit (t, beta, gamma; ref where the ray hits), through
the C ABI with the default persistent launch shape."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd import scenes
from rta_amd.abi import HIT_DTYPE, NODE_DTYPE, TRI_DTYPE, call, ptr
from rta_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
F32MAX = np.float32(np.finfo(np.float32).max)


def say(msg):
    print(f"[full-size] {msg}", flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def records_equal(got, exp, what):
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(got[f]), bits(exp[f])), f"{what}: {f}"
    hit = exp["t"] != F32MAX
    assert np.array_equal(got["ref"][hit], exp["ref"][hit]), f"{what}: ref"
    return hit


def test_terrain_benchmark_workload_bit_exact(gpu):
    say("scene + oracle build")
    tris_h = scenes.terrain_scene()
    n = len(tris_h)
    cam = scenes.terrain_camera()
    R = cam.w * cam.h
    ob = oracle.lbvh_build(tris_h)

    say("GPU build + pack")
    tris = DeviceBuffer.from_array(tris_h)
    nodes = DeviceBuffer(32 * (2 * n - 1))
    idx = DeviceBuffer(4 * n)
    status = DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", n, C.byref(sz))
    ws = DeviceBuffer(sz.value)
    call("rta_lbvh_build", ptr(tris), n, ptr(nodes), ptr(idx), ptr(ws), C.c_size_t(sz.value),
         ptr(status), None, None)
    pairs, tri4, flags = DeviceBuffer(64 * (2 * n - 1)), DeviceBuffer(48 * n), DeviceBuffer(4)
    call("rta_bvh_pack", ptr(pairs), ptr(tri4), ptr(flags), ptr(nodes), 2 * n - 1, ptr(tris), n,
         None)
    rta_amd.synchronize()
    got_nodes = nodes.download(NODE_DTYPE, 2 * n - 1)
    for f in ("type_left_elems", "right_tris", "min", "max"):
        assert np.array_equal(got_nodes[f], ob["nodes"][f]), f
    assert np.array_equal(bits(tris.download(TRI_DTYPE, n)), bits(ob["tris"]))
    assert int(flags.download(np.uint32, 1)[0]) == 0      # the exact-division path
    root = (int(got_nodes[0]["type_left_elems"]), int(got_nodes[0]["right_tris"]))

    say("packed build (the layout written by the build's own passes)")
    t2 = DeviceBuffer.from_array(tris_h)
    nodes2, idx2 = DeviceBuffer(32 * (2 * n - 1)), DeviceBuffer(4 * n)
    pairs2, tri42, flags2 = DeviceBuffer(64 * (2 * n - 1)), DeviceBuffer(48 * n), DeviceBuffer(4)
    call("rta_lbvh_build_packed", ptr(t2), n, ptr(nodes2), ptr(idx2), ptr(pairs2), ptr(tri42),
         ptr(flags2), ptr(ws), C.c_size_t(sz.value), ptr(status), None, None)
    rta_amd.synchronize()
    assert nodes2.download(NODE_DTYPE, 2 * n - 1).tobytes() == got_nodes.tobytes()
    assert np.array_equal(idx2.download(np.uint32, n), idx.download(np.uint32, n))
    assert np.array_equal(tri42.download(np.uint8, 48 * n), tri4.download(np.uint8, 48 * n))
    inner = (got_nodes["type_left_elems"] & 1) == 1
    assert np.array_equal(pairs2.download(np.uint8, 64 * (2 * n - 1)).reshape(-1, 64)[inner],
                          pairs.download(np.uint8, 64 * (2 * n - 1)).reshape(-1, 64)[inner])
    assert int(flags2.download(np.uint32, 1)[0]) == 0
    del t2, nodes2, idx2, pairs2, tri42, flags2

    say("primary rays")
    prim = rta_amd.RayBuffers(R)
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    call("rta_setup_shirley_rays", ptr(prim.gpu_direction), ptr(prim.gpu_origin),
         ptr(prim.gpu_maxt), cam.fovy, cam.aspect, cam.w, cam.h, f3(cam.view_dir), f3(cam.pos),
         f3(cam.up), float(F32MAX), None)
    o, d, t = cam.host_rays()
    go, gd, gt = prim.download(R)
    assert np.array_equal(bits(go), bits(o)) and np.array_equal(bits(gd), bits(d))
    hits_p, err, queue = DeviceBuffer(16 * R), DeviceBuffer(4), DeviceBuffer(4)
    err.fill_u32(0)
    call("rta_reset_intersections", ptr(hits_p), R, 1, None)
    call("rta_trace_cis_persistent", ptr(tri4), ptr(pairs), ptr(flags), root[0], root[1],
         ptr(prim.gpu_origin), ptr(prim.gpu_direction), ptr(prim.gpu_maxt), R, None,
         ptr(hits_p), ptr(err), ptr(queue), None, None)
    rta_amd.synchronize()
    ref_p, deep = oracle.trace_kernel(ob["tris"], ob["nodes"], o, d, t)
    assert deep == 0 and int(err.download(np.uint32, 1)[0]) == 0
    hit = records_equal(hits_p.download(HIT_DTYPE, R), ref_p, "primary")
    assert hit.mean() > 0.9

    say("bounce rays")
    bnc = rta_amd.RayBuffers(R)
    src, count = DeviceBuffer(4 * R), DeviceBuffer(4)
    call("rta_bounce_work_bytes", R, C.byref(sz))
    bwork = DeviceBuffer(sz.value)
    call("rta_diffuse_bounce", ptr(bnc.gpu_origin), ptr(bnc.gpu_direction), ptr(bnc.gpu_maxt),
         ptr(src), ptr(count), ptr(hits_p), ptr(tris), ptr(prim.gpu_origin),
         ptr(prim.gpu_direction), R, C.c_uint64(scenes.BOUNCE_SEED), ptr(bwork),
         C.c_size_t(sz.value), None)
    rta_amd.synchronize()
    bo, bd, bt, bsrc = oracle.diffuse_bounce(ref_p, ob["tris"], o, d, scenes.BOUNCE_SEED)
    nb = int(count.download(np.uint32, 1)[0])
    assert nb == len(bt) == 2_064_937
    go, gd, gt = bnc.download(nb)
    assert np.array_equal(src.download(np.uint32, nb), bsrc)
    assert np.array_equal(bits(go), bits(bo)) and np.array_equal(bits(gd), bits(bd))
    assert np.array_equal(bits(gt), bits(bt))

    say("bounce trace")
    hits_b = DeviceBuffer(16 * R)
    call("rta_reset_intersections", ptr(hits_b), R, 1, None)
    call("rta_trace_cis_persistent", ptr(tri4), ptr(pairs), ptr(flags), root[0], root[1],
         ptr(bnc.gpu_origin), ptr(bnc.gpu_direction), ptr(bnc.gpu_maxt), R, ptr(count),
         ptr(hits_b), ptr(err), ptr(queue), None, None)
    rta_amd.synchronize()
    ref_b, deep = oracle.trace_kernel(ob["tris"], ob["nodes"], bo, bd, bt)
    assert deep == 0 and int(err.download(np.uint32, 1)[0]) == 0
    hit = records_equal(hits_b.download(HIT_DTYPE, nb), ref_b, "bounce")
    assert 0.05 < hit.mean() < 0.95
    say("done")


def test_batch_workload_full_size_bit_exact(gpu):
    """The 256-mesh workload (7,031,454 triangles, 15,327,936 rays): every
    mesh's tree, permutation, reordered triangles and hits from the batched
    launches equal the oracle's build and trace of that mesh alone."""
    from rta_amd.batch import MeshBatch

    say("batch: generate")
    meshes, rays = scenes.batch_meshes(256)
    assert sum(len(m) for m in meshes) == 7_031_454
    assert sum(len(r[2]) for r in rays) == 15_327_936
    b = MeshBatch(meshes, rays)
    say("batch: build + trace")
    b.build()
    b.trace()
    b.stream.synchronize()
    b.check_errors()
    assert int(b.status.download(np.uint32, 1)[0]) & ~1 == 0
    say("batch: compare")
    for m, (mesh, (o, d, t)) in enumerate(zip(meshes, rays)):
        got = b.mesh_results(m)
        exp = oracle.lbvh_build(mesh)
        assert np.array_equal(got["indices"], exp["indices"]), f"mesh {m} indices"
        assert np.array_equal(bits(got["tris"]), bits(exp["tris"])), f"mesh {m} tris"
        for f in ("type_left_elems", "right_tris", "min", "max"):
            assert np.array_equal(got["nodes"][f], exp["nodes"][f]), f"mesh {m} nodes.{f}"
        ref, deep = oracle.trace_kernel(exp["tris"], exp["nodes"], o, d, t)
        assert deep == 0
        g, hit = got["hits"], ref["t"] != F32MAX
        assert np.array_equal(g["t"] != F32MAX, hit), f"mesh {m} hit flags"
        for f in ("t", "beta", "gamma"):
            assert np.array_equal(bits(g[f][hit]), bits(ref[f][hit])), f"mesh {m} {f}"
        assert np.array_equal(g["ref"][hit], ref["ref"][hit]), f"mesh {m} ref"
    say("batch: done")


def test_dynamic_frame_full_size_bit_exact(gpu):
    """One eager frame of the 8,388,608-triangle dynamic scene at 3840x2160
    (jitter, rebuild, primary rays, two diffuse bounces, a shadow ray per
    hit of each level) against the oracle's chain of the same frame."""
    from rta_amd.frame import DynamicFrame
    from test_dynamic import SEED, _compare, _oracle_frame

    say("dynamic: generate")
    v, i = scenes.instances(64, 256, seed=scenes.DYNAMIC_SEED)
    assert len(i) // 3 == 8_388_608
    cam = scenes.dynamic_camera(3840, 2160, n_inst=64)
    light = scenes.dynamic_light(64)
    fr = DynamicFrame(v, i, cam, light, scenes.DYNAMIC_SIGMA, SEED, bounces=2, shadows=True)
    fr.prime()
    say("dynamic: frame 1")
    fr.enqueue(frame=1)
    fr.stream.synchronize()
    fr.check_errors()
    res = fr.results()
    say("dynamic: oracle")
    exp = _oracle_frame(v, i, cam, light, 1)
    _compare(res, exp)
    say("dynamic: done")
