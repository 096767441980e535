"""Batches of many independent meshes, each with its own rays: all LBVH
builds in one launch per stage and all traversals in one persistent launch
(rta_batch_* in include/rta_hip.h), with an autotuned launch shape."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import HIT_DTYPE, NODE_DTYPE, TRI_DTYPE, TraceConfig, call, ptr
from .device import DeviceBuffer, Stream


class MeshBatch:
    """Device buffers of a batch.  meshes: list of TRI_DTYPE arrays; rays:
    list of (origins[3r], directions[3r], max_t[r]) per mesh (may be empty)."""

    def __init__(self, meshes, rays=None, stream: Stream | None = None):
        self.M = len(meshes)
        if rays is None:
            rays = [(np.zeros(0, np.float32),) * 3] * self.M
        self.n_tris = np.array([len(m) for m in meshes], np.int32)
        self.n_rays = np.array([len(r[2]) for r in rays], np.int32)
        h = C.c_void_p()
        call("rta_batch_create", self.M, ptr(self.n_tris), ptr(self.n_rays), C.byref(h))
        self.handle = h
        tt, tn, tr = C.c_longlong(), C.c_longlong(), C.c_longlong()
        call("rta_batch_sizes", h, C.byref(tt), C.byref(tn), C.byref(tr))
        self.total_tris, self.total_nodes, self.total_rays = tt.value, tn.value, tr.value
        self.tri_off = np.zeros(self.M, np.uint32)
        self.node_off = np.zeros(self.M, np.uint32)
        self.ray_off = np.zeros(self.M, np.uint32)
        call("rta_batch_offsets", h, ptr(self.tri_off), ptr(self.node_off), ptr(self.ray_off))
        self.stream = stream if stream is not None else Stream()
        allt = np.concatenate(meshes)
        self.pristine = DeviceBuffer.from_array(allt)
        self.tris = DeviceBuffer(104 * self.total_tris)
        self.nodes = DeviceBuffer(32 * self.total_nodes)
        self.indices = DeviceBuffer(4 * self.total_tris)
        self.status = DeviceBuffer(4)
        sz = C.c_size_t(0)
        call("rta_batch_workspace_bytes", h, C.byref(sz))
        self.ws, self.ws_bytes = DeviceBuffer(sz.value), sz.value
        self.pairs = DeviceBuffer(64 * self.total_nodes)
        self.tri4 = DeviceBuffer(48 * self.total_tris)
        self.flags = DeviceBuffer(4 * self.M)
        R = max(1, self.total_rays)
        if self.total_rays:
            self.ro = DeviceBuffer.from_array(np.concatenate([r[0] for r in rays]).astype(np.float32))
            self.rd = DeviceBuffer.from_array(np.concatenate([r[1] for r in rays]).astype(np.float32))
            self.rt = DeviceBuffer.from_array(np.concatenate([r[2] for r in rays]).astype(np.float32))
        self.hits = DeviceBuffer(16 * R)
        self.err = DeviceBuffer(4)
        self.err.fill_u32(0)
        self.queue = DeviceBuffer(4)
        self.config = TraceConfig(64, 16, 32, 0, 10)
        # uploads and fills above ran on the default stream
        call("rta_hip_device_synchronize")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                call("rta_batch_destroy", h)
            except Exception:
                pass
            self.handle = None

    def build(self, packed: bool = True) -> None:
        """Restore the input triangles and build every mesh with its
        traversal layout: in the build's own passes (rta_batch_build_packed),
        or with packed=False as a build followed by rta_batch_pack."""
        s = ptr(self.stream)
        call("rta_hip_memcpy_dtod", ptr(self.tris), ptr(self.pristine),
             C.c_size_t(104 * self.total_tris), s)
        if packed:
            call("rta_batch_build_packed", self.handle, ptr(self.tris), ptr(self.nodes),
                 ptr(self.indices), ptr(self.pairs), ptr(self.tri4), ptr(self.flags),
                 ptr(self.ws), C.c_size_t(self.ws_bytes), ptr(self.status), s)
            return
        call("rta_batch_build", self.handle, ptr(self.tris), ptr(self.nodes), ptr(self.indices),
             ptr(self.ws), C.c_size_t(self.ws_bytes), ptr(self.status), s)
        call("rta_batch_pack", self.handle, ptr(self.pairs), ptr(self.tri4), ptr(self.flags),
             ptr(self.nodes), ptr(self.tris), s)

    def reset_hits(self) -> None:
        call("rta_reset_intersections", ptr(self.hits), max(1, self.total_rays), 1,
             ptr(self.stream))

    def trace(self, any_hit: bool = False, config: TraceConfig | None = None) -> None:
        self.reset_hits()
        cfg = config if config is not None else self.config
        call("rta_batch_trace_persistent", self.handle, ptr(self.tri4), ptr(self.pairs),
             ptr(self.flags), ptr(self.ro), ptr(self.rd), ptr(self.rt), ptr(self.hits),
             ptr(self.err), ptr(self.queue), int(bool(any_hit)), C.byref(cfg), ptr(self.stream))

    def autotune(self, any_hit: bool = False, rounds: int = 2, reps: int = 2):
        """Pick the fastest launch shape on this batch; returns (config, ms)."""
        self.reset_hits()
        best = TraceConfig(*self.config.as_tuple())
        ms = C.c_float(0)
        call("rta_batch_autotune", self.handle, ptr(self.tri4), ptr(self.pairs), ptr(self.flags),
             ptr(self.ro), ptr(self.rd), ptr(self.rt), ptr(self.hits), ptr(self.err),
             ptr(self.queue), int(bool(any_hit)), rounds, reps, C.byref(best), C.byref(ms),
             ptr(self.stream))
        self.config = best
        return best, ms.value

    def capture(self, any_hit: bool = False, config: TraceConfig | None = None):
        """The whole batch frame -- restore the triangles, build every mesh
        with its traversal layout, reset the hits, trace all rays -- captured
        as one graph
        (rta_hip_graph_*), replayed with .launch(self.stream)."""
        from .frame import Graph

        call("rta_hip_graph_begin", ptr(self.stream))
        h = C.c_void_p()
        try:
            self.build()
            self.trace(any_hit=any_hit, config=config)
        finally:
            # always end the capture so the stream is usable again
            call("rta_hip_graph_end", ptr(self.stream), C.byref(h))
        return Graph(h)

    def overflow_count(self) -> int:
        """Rays the last traversal with stack 16 or 24 traced again with the
        full 32 entries."""
        c = C.c_uint32(0)
        call("rta_batch_overflow_count", self.handle, C.byref(c), ptr(self.stream))
        return int(c.value)

    def check_errors(self) -> None:
        e = int(self.err.download(np.uint32, 1)[0])
        if e:
            raise RuntimeError(f"{e} rays overflowed the traversal stack")

    def mesh_results(self, m: int) -> dict:
        """Host copies of mesh m's build and trace outputs (local indices)."""
        self.stream.synchronize()
        n, t0, n0 = int(self.n_tris[m]), int(self.tri_off[m]), int(self.node_off[m])
        out = {
            "tris": self.tris.download(TRI_DTYPE, n, offset=104 * t0),
            "nodes": self.nodes.download(NODE_DTYPE, 2 * n - 1, offset=32 * n0),
            "indices": self.indices.download(np.uint32, n, offset=4 * t0),
        }
        r, r0 = int(self.n_rays[m]), int(self.ray_off[m])
        out["hits"] = self.hits.download(HIT_DTYPE, r, offset=16 * r0) if r else None
        return out
