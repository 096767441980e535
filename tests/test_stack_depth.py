"""Traversal stack at and beyond its 32 entries (bbvh-cuda/kernels.cu:787).

A hand-made chain tree makes every ray along +x push one leaf per level:
inner node I_k has the leaf L_k (a triangle far along x) as left child and
I_{k+1} (a box near the ray origin) as right child, so the right child is
entered and L_k pushed.  After `pushes` levels the chain ends at an inner
node whose right child lies off the ray: that step keeps one child and
pushes nothing.  The closest hit is the leaf pushed last (L_{pushes-1}, at
x = 120; every other leaf lies at x >= 150), i.e. the stack's top entry at
the deepest point.

With 32 pushes a ray holds exactly 32 entries (the reference's limit, which
it does not check); every kernel must then return the oracle's record.
With 33 pushes the ray needs a 33rd entry: every kernel must count a stack
overflow (RTA_E_STACK) and return normally.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd.abi import HIT_DTYPE, NODE_DTYPE, TRI_DTYPE, RTA_E_STACK, RtaError, TraceConfig
from rta_amd.device import DeviceBuffer

F32MAX = np.float32(np.finfo(np.float32).max)


def say(msg):
    print(f"[stack-depth] {msg}", flush=True)
    sys.stdout.flush()


def chain_tree(pushes: int):
    """(triangles, nodes) of the chain described above.  Inner nodes
    I_0..I_{pushes+1} take indices 0..pushes+1, leaves L_0..L_{pushes+2}
    the indices after them; leaf L_k holds triangle k."""
    n_inner = pushes + 2
    n = pushes + 3                       # one triangle per leaf
    tris = np.zeros(n, TRI_DTYPE)
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    leaf = lambda k: n_inner + k         # noqa: E731
    xs = np.array([150.0 + k for k in range(n)], np.float32)
    xs[pushes - 1] = 120.0               # the closest hit: the last pushed leaf
    for k in range(n):
        if k < pushes + 1:               # on the ray's path (y, z in [-1, 3])
            y0, z0 = -1.0, -1.0
        else:                            # the off-ray subtree's leaves
            y0, z0 = 5.0, 5.0
        tris[k]["a"] = (xs[k], y0, z0)
        tris[k]["b"] = (xs[k], y0 + 4.0, z0)
        tris[k]["c"] = (xs[k], y0, z0 + 4.0)
        nd = nodes[leaf(k)]
        nd["type_left_elems"] = 1 << 1   # a leaf of one triangle
        nd["right_tris"] = k
        nd["min"] = np.minimum(np.minimum(tris[k]["a"], tris[k]["b"]), tris[k]["c"])
        nd["max"] = np.maximum(np.maximum(tris[k]["a"], tris[k]["b"]), tris[k]["c"])
    for k in range(n_inner):
        nd = nodes[k]
        if k <= pushes:
            nd["type_left_elems"] = (leaf(k) << 1) | 1
            nd["right_tris"] = k + 1
        else:                            # I_{pushes+1}: two off-ray leaves
            nd["type_left_elems"] = (leaf(k) << 1) | 1
            nd["right_tris"] = leaf(k + 1)
        if k == 0:
            nd["min"], nd["max"] = (0.0, -1.0, -1.0), (200.0, 9.0, 9.0)
        elif k <= pushes:                # near the ray origin: entered first
            nd["min"], nd["max"] = (float(k), -1.0, -1.0), (100.0, 3.0, 3.0)
        else:                            # off the ray
            nd["min"], nd["max"] = (float(k), 5.0, 5.0), (100.0, 9.0, 9.0)
    return tris, nodes


def chain_rays(n_random: int = 301, seed: int = 4):
    """Rays along +x through the chain: one with small nonzero y/z slopes (the
    branch-free exact-division slab path), one axis-parallel (the IEEE path),
    and random ones from x = -1 with small slopes."""
    rng = np.random.default_rng(seed)
    o = [(-1.0, 0.01, 0.03), (-1.0, 0.2, 0.3)]
    d = [(1.0, 1e-3, 2e-3), (1.0, 0.0, 0.0)]
    ro = np.concatenate([np.array(o, np.float32),
                         np.column_stack([np.full(n_random, -1.0),
                                          rng.uniform(-0.5, 0.5, (n_random, 2))])])
    rd = np.concatenate([np.array(d, np.float32),
                         np.column_stack([np.ones(n_random),
                                          rng.uniform(-2e-3, 2e-3, (n_random, 2))])])
    rd = rd / np.linalg.norm(rd, axis=1, keepdims=True)
    ro, rd = ro.astype(np.float32), rd.astype(np.float32)
    return ro.reshape(-1), rd.reshape(-1), np.full(len(ro), F32MAX, np.float32)


# ------------------------------------------------------------- CPU checks

def test_chain_depth_on_the_oracle():
    """The oracle holds exactly 32 entries for 32 pushes and needs 33 for
    33 pushes; the closest hit is the last pushed leaf (triangle pushes-1)."""
    for pushes, depth_exp, deep_exp in ((32, 32, 0), (33, 33, None)):
        tris, nodes = chain_tree(pushes)
        o, d, t = chain_rays()
        hits, deep, depth = oracle.trace_kernel(tris, nodes, o, d, t, depth=True)
        assert (depth == depth_exp).all(), (pushes, np.unique(depth))
        assert (hits["t"] != F32MAX).all()
        assert (hits["ref"] == pushes - 1).all()
        assert deep == (len(t) if deep_exp is None else deep_exp)
        # x = 120 is reached at t = 121 along the axis-parallel ray
        assert hits["t"][1] == np.float32(121.0)


def test_chain_oracle_agrees_with_cpu_plugin_tracer():
    """The CPU plugin's child-is tracer (128-entry stack) gives the same
    records on the 32-entry chain: the kernel semantics lose nothing there."""
    tris, nodes = chain_tree(32)
    o, d, t = chain_rays()
    k, _ = oracle.trace_kernel(tris, nodes, o, d, t)
    c, _, _ = oracle.trace_cis(tris, nodes, o, d, t, len(t))
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(k[f].view(np.uint32), c[f].view(np.uint32)), f


def test_chain_plants_a_lost_entry_in_the_oracle():
    """The mistake the persistent kernel once made, planted in a copy of the
    chain on the CPU: replacing the top pushed leaf by the culled sibling
    (what a clobbered top slot holds) changes the closest hit, so the GPU
    test below detects a lost entry."""
    tris, nodes = chain_tree(32)
    o, d, t = chain_rays()
    good, _ = oracle.trace_kernel(tris, nodes, o, d, t)
    broken = nodes.copy()
    broken[31]["type_left_elems"] = (33 << 1) | 1   # L_31 replaced by off-ray I_33
    bad, _ = oracle.trace_kernel(tris, broken, o, d, t)
    assert not np.array_equal(good["ref"], bad["ref"])


# ------------------------------------------------------------- GPU checks

def _trace(tris, nodes, o, d, t, layout, config=None):
    n = len(tris)
    bvh = rta_amd.BinaryLbvh(DeviceBuffer.from_array(tris), n, DeviceBuffer.from_array(nodes),
                             DeviceBuffer(4 * n))
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    bouncer = rta_amd.GpuBouncer(len(t))
    tracer = rta_amd.CisTracer(rays, bvh, bouncer, layout=layout, config=config)
    tracer.prepare_trace()
    tracer.launch()
    rta_amd.synchronize()
    return tracer, bouncer


LAYOUT_CASES = [("plain", None), ("packed", None), ("persistent", None),
                ("persistent", (128, 16, 7, 0, 3)), ("persistent", (256, 64, 1, 0, 64))]


@pytest.mark.gpu
@pytest.mark.parametrize("layout,shape", LAYOUT_CASES)
def test_exactly_32_entries(gpu, layout, shape):
    say(f"32 entries layout={layout} shape={shape}")
    tris, nodes = chain_tree(32)
    o, d, t = chain_rays()
    ref, deep = oracle.trace_kernel(tris, nodes, o, d, t)
    assert deep == 0
    cfg = TraceConfig(*shape) if shape else None
    tracer, bouncer = _trace(tris, nodes, o, d, t, layout, cfg)
    tracer.check_errors()
    got = bouncer.intersections()
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(got[f].view(np.uint32), ref[f].view(np.uint32)), f


@pytest.mark.gpu
@pytest.mark.parametrize("layout,shape", LAYOUT_CASES)
def test_33_entries_report_stack_overflow(gpu, layout, shape):
    say(f"33 entries layout={layout} shape={shape}")
    tris, nodes = chain_tree(33)
    o, d, t = chain_rays()
    _, deep = oracle.trace_kernel(tris, nodes, o, d, t)
    assert deep == len(t)
    cfg = TraceConfig(*shape) if shape else None
    tracer, _ = _trace(tris, nodes, o, d, t, layout, cfg)
    with pytest.raises(RtaError) as e:
        tracer.check_errors()
    assert e.value.code == RTA_E_STACK
    # the device is still usable: the 32-entry chain traces correctly after it
    tris, nodes = chain_tree(32)
    tracer, bouncer = _trace(tris, nodes, o, d, t, layout, cfg)
    tracer.check_errors()
    assert (bouncer.intersections()["ref"] == 31).all()
