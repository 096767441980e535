"""Single-mesh persistent traversal with a short LDS stack
(rta_trace_cis_persistent_ovf, rta_trace_shadow_cis_persistent_ovf): with
16 or 24 entries a ray that needs more is put on the overflow list and
traced again with the reference's 32 (bbvh-cuda/kernels.cu:787), so every
record equals the 32-entry traversal's.  Checked on the hand-made chain
trees of test_stack_depth (every ray needs exactly `pushes` entries, so the
list holds all rays or none) against the oracle, and on a heightfield +
random scene's primary and bounce rays against the 32-entry launch."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd import scenes
from rta_amd.abi import (HIT_DTYPE, NODE_DTYPE, RTA_E_INVALID, TraceConfig, call, ptr)
from rta_amd.device import DeviceBuffer

from test_stack_depth import chain_rays, chain_tree

F32MAX = np.float32(np.finfo(np.float32).max)


def say(msg):
    print(f"[short-stack] {msg}", flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def same_records(got, exp):
    """t, beta, gamma everywhere (a miss keeps the reset sentinel), ref
    where the ray hits."""
    for f in ("t", "beta", "gamma"):
        if not np.array_equal(bits(got[f]), bits(exp[f])):
            return False
    hit = exp["t"] != F32MAX
    return np.array_equal(got["ref"][hit], exp["ref"][hit])


class Packed:
    """Packed layout of a node array (rta_bvh_pack) and a persistent launch
    over it, with or without the overflow list."""

    def __init__(self, tris, nodes):
        n = len(tris)
        self.n = n
        self.tris = DeviceBuffer.from_array(tris)
        dn = DeviceBuffer.from_array(nodes)
        self.pairs = DeviceBuffer(64 * (2 * n - 1))
        self.tri4 = DeviceBuffer(48 * n)
        self.flags = DeviceBuffer(4)
        call("rta_bvh_pack", ptr(self.pairs), ptr(self.tri4), ptr(self.flags), ptr(dn), 2 * n - 1,
             ptr(self.tris), n, None)
        self.root = (int(nodes[0]["type_left_elems"]), int(nodes[0]["right_tris"]))
        self.err, self.queue = DeviceBuffer(4), DeviceBuffer(4)

    def trace(self, ro, rd, rt, n_rays, stack, shadow=False, count=None, ovf=None, block=64):
        hits = DeviceBuffer(16 * n_rays)
        call("rta_reset_intersections", ptr(hits), n_rays, 1, None)
        self.err.fill_u32(0)
        cfg = TraceConfig(block, 16, 32, 0, 10, stack)
        args = (ptr(self.tri4), ptr(self.pairs), ptr(self.flags), self.root[0], self.root[1],
                ptr(ro), ptr(rd), ptr(rt), n_rays, ptr(count), ptr(hits), ptr(self.err),
                ptr(self.queue), C.byref(cfg))
        name = "rta_trace_shadow_cis_persistent" if shadow else "rta_trace_cis_persistent"
        if ovf is None:
            call(name, *args, None)
        else:
            call(name + "_ovf", *args, ptr(ovf), None)
        rta_amd.synchronize()
        assert int(self.err.download(np.uint32, 1)[0]) == 0
        return hits.download(HIT_DTYPE, n_rays)


@pytest.mark.gpu
@pytest.mark.parametrize("pushes", [14, 20, 32])
def test_chain_rays_overflow_and_match_the_oracle(gpu, pushes):
    """Every ray of a `pushes`-entry chain overflows a stack of fewer entries
    (the list then holds every ray) and none a deeper one; the records equal
    the oracle's either way, closest hit and any hit."""
    tris, nodes = chain_tree(pushes)
    o, d, t = chain_rays()
    ref, deep = oracle.trace_kernel(tris, nodes, o, d, t)
    assert deep == 0
    m = len(t)
    pk = Packed(tris, nodes)
    ro, rd, rt = (DeviceBuffer.from_array(a) for a in (o, d, t))
    ovf = DeviceBuffer(4 * (m + 1))
    for stack in (16, 24, 32):
        for block in (64, 128):
            got = pk.trace(ro, rd, rt, m, stack, ovf=ovf, block=block)
            listed = int(ovf.download(np.uint32, 1)[0]) if stack < 32 else 0
            say(f"pushes {pushes} stack {stack} block {block}: {listed} of {m} rays retraced")
            assert listed == (m if pushes > stack else 0)
            if listed:
                assert sorted(ovf.download(np.uint32, m + 1)[1:]) == list(range(m))
            assert same_records(got, ref), (stack, block)
        shadow = pk.trace(ro, rd, rt, m, stack, shadow=True, ovf=ovf)
        assert (shadow["t"] != F32MAX).all()        # every chain ray hits something


def _scene_rays():
    """A 128 x 128 heightfield with 4,000 random triangles, 320 x 180
    primary rays and the GPU's diffuse bounce of every hit."""
    tris_h = scenes.terrain_scene(grid=128, n_random=4000, seed=11)
    n = len(tris_h)
    tris = DeviceBuffer.from_array(tris_h)
    nodes, idx, status = DeviceBuffer(32 * (2 * n - 1)), DeviceBuffer(4 * n), DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", n, C.byref(sz))
    ws = DeviceBuffer(sz.value)
    call("rta_lbvh_build", ptr(tris), n, ptr(nodes), ptr(idx), ptr(ws), C.c_size_t(sz.value),
         ptr(status), None, None)
    rta_amd.synchronize()
    nodes_h = nodes.download(NODE_DTYPE, 2 * n - 1)
    tris_sorted = tris.download(rta_amd.abi.TRI_DTYPE, n)
    pk = Packed(tris_sorted, nodes_h)
    cam = scenes.terrain_camera(320, 180)
    R = cam.w * cam.h
    prim = rta_amd.RayBuffers(R)
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])  # noqa: E731
    call("rta_setup_shirley_rays", ptr(prim.gpu_direction), ptr(prim.gpu_origin),
         ptr(prim.gpu_maxt), cam.fovy, cam.aspect, cam.w, cam.h, f3(cam.view_dir), f3(cam.pos),
         f3(cam.up), float(F32MAX), None)
    return pk, tris, prim, R


@pytest.mark.gpu
def test_scene_primary_and_bounce_rays_equal_full_stack(gpu):
    pk, tris, prim, R = _scene_rays()
    ovf = DeviceBuffer(4 * (R + 1))
    p32 = pk.trace(prim.gpu_origin, prim.gpu_direction, prim.gpu_maxt, R, 32)
    hp = DeviceBuffer.from_array(p32)
    bnc = rta_amd.RayBuffers(R)
    src, count = DeviceBuffer(4 * R), DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_bounce_work_bytes", R, C.byref(sz))
    bwork = DeviceBuffer(sz.value)
    call("rta_diffuse_bounce", ptr(bnc.gpu_origin), ptr(bnc.gpu_direction), ptr(bnc.gpu_maxt),
         ptr(src), ptr(count), ptr(hp), ptr(tris), ptr(prim.gpu_origin), ptr(prim.gpu_direction),
         R, C.c_uint64(scenes.BOUNCE_SEED), ptr(bwork), C.c_size_t(sz.value), None)
    rta_amd.synchronize()
    nb = int(count.download(np.uint32, 1)[0])
    assert 0 < nb <= R
    for name, rays, cnt, m in (("primary", prim, None, R), ("bounce", bnc, count, nb)):
        full = pk.trace(rays.gpu_origin, rays.gpu_direction, rays.gpu_maxt, R, 32, count=cnt)[:m]
        full_s = pk.trace(rays.gpu_origin, rays.gpu_direction, rays.gpu_maxt, R, 32, shadow=True,
                          count=cnt)[:m]
        for stack in (16, 24):
            got = pk.trace(rays.gpu_origin, rays.gpu_direction, rays.gpu_maxt, R, stack,
                           count=cnt, ovf=ovf)[:m]
            listed = int(ovf.download(np.uint32, 1)[0])
            assert listed <= m
            got_s = pk.trace(rays.gpu_origin, rays.gpu_direction, rays.gpu_maxt, R, stack,
                             shadow=True, count=cnt, ovf=ovf)[:m]
            say(f"{name} ({m} rays) stack {stack}: {listed} retraced (closest hit)")
            assert same_records(got, full), (name, stack)
            assert same_records(got_s, full_s), (name, stack, "shadow")


def test_ovf_entry_points_check_their_arguments():
    """Checked before any device work: a NULL overflow list, and stack 20."""
    lib = rta_amd.hip()
    fake = C.c_void_p(64)
    for name in ("rta_trace_cis_persistent_ovf", "rta_trace_shadow_cis_persistent_ovf"):
        fn = getattr(lib, name)
        cfg = TraceConfig(64, 16, 32, 0, 10, 16)
        rc = fn(fake, fake, fake, 1, 0, fake, fake, fake, 10, None, fake, fake, fake,
                C.byref(cfg), None, None)
        assert rc == RTA_E_INVALID and b"overflow" in lib.rta_hip_last_error()
        cfg = TraceConfig(64, 16, 32, 0, 10, 20)
        rc = fn(fake, fake, fake, 1, 0, fake, fake, fake, 10, None, fake, fake, fake,
                C.byref(cfg), fake, None)
        assert rc == RTA_E_INVALID and b"stack" in lib.rta_hip_last_error()
