"""rta_lbvh_build_noalloc: the LBVH build with its workspace in the
library's fixed device array, against rta_lbvh_build with a caller
workspace (the entry point it varies).  Outputs must be bit-identical:
nodes, the index buffer, the reordered triangles and the status word.
Two calls in a row, the second with a smaller input, so a call that read
what the previous one left in the shared array would show."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import rta_amd
from rta_amd import scenes
from rta_amd.abi import (NODE_DTYPE, RTA_E_INVALID, RTA_E_WORKSPACE, RTA_LBVH_NOALLOC_MAX_TRIS,
                         TRI_DTYPE, RtaError, call, ptr)
from rta_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def _with_workspace(tris):
    n = len(tris)
    t = DeviceBuffer.from_array(tris)
    nodes, idx, st = DeviceBuffer(32 * (2 * n - 1)), DeviceBuffer(4 * n), DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", n, C.byref(sz))
    ws = DeviceBuffer(sz.value)
    call("rta_lbvh_build", ptr(t), n, ptr(nodes), ptr(idx), ptr(ws), C.c_size_t(sz.value),
         ptr(st), None, None)
    rta_amd.synchronize()
    return (nodes.download(NODE_DTYPE, 2 * n - 1), idx.download(np.uint32, n),
            t.download(TRI_DTYPE, n), int(st.download(np.uint32, 1)[0]))


def _scenes():
    big = np.concatenate([scenes.heightfield(320, 320, seed=21),          # 204,800
                          scenes.random_triangles(57_000, seed=22)])      # 261,800 in all
    small = np.concatenate([scenes.icosphere(3, jitter=0.02, seed=5),
                            scenes.random_triangles(8_727, seed=23)])     # 10,007
    assert len(big) <= RTA_LBVH_NOALLOC_MAX_TRIS and len(small) < len(big)
    return big, small


def test_noalloc_build_equals_workspace_build(gpu):
    big, small = _scenes()
    refs = [_with_workspace(x) for x in (big, small)]
    # the outputs and inputs of both calls exist before the first one runs,
    # so the two no-allocation builds run back to back
    bufs = []
    for x in (big, small):
        n = len(x)
        bufs.append((DeviceBuffer.from_array(x), DeviceBuffer(32 * (2 * n - 1)),
                     DeviceBuffer(4 * n), DeviceBuffer(4)))
    for x, (t, nodes, idx, st) in zip((big, small), bufs):
        say(f"rta_lbvh_build_noalloc n={len(x)}")
        call("rta_lbvh_build_noalloc", ptr(t), len(x), ptr(nodes), ptr(idx), ptr(st), None)
    rta_amd.synchronize()
    for x, (t, nodes, idx, st), (r_nodes, r_idx, r_tris, r_st) in zip((big, small), bufs, refs):
        n = len(x)
        got_nodes = nodes.download(NODE_DTYPE, 2 * n - 1)
        assert got_nodes.tobytes() == r_nodes.tobytes()
        assert np.array_equal(idx.download(np.uint32, n), r_idx)
        assert t.download(TRI_DTYPE, n).tobytes() == r_tris.tobytes()
        assert int(st.download(np.uint32, 1)[0]) == r_st
        say(f"n={n}: nodes, indices, triangles and status equal the workspace build")


def test_noalloc_build_rejects_sizes_beyond_its_limit(gpu):
    n = RTA_LBVH_NOALLOC_MAX_TRIS + 1
    with pytest.raises(RtaError) as e:
        # pointers are never touched: the size check comes first
        call("rta_lbvh_build_noalloc", C.c_void_p(16), n, C.c_void_p(16), C.c_void_p(16), None,
             None)
    assert e.value.code == RTA_E_WORKSPACE
    assert "RTA_LBVH_NOALLOC_MAX_TRIS" in str(e.value)
    with pytest.raises(RtaError) as e:
        call("rta_lbvh_build_noalloc", C.c_void_p(16), 0, C.c_void_p(16), C.c_void_p(16), None,
             None)
    assert e.value.code == RTA_E_INVALID
