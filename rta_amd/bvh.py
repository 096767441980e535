"""Python host mirror of the reference's GPU BVH plugin objects.

Names and roles follow the reference so that tests read like its interface:

  LbvhConstructor   lbvh_constructor          bbvh-cuda/lbvh.h:49-216
  BinaryLbvh        binary_lbvh/cuda::binary_bvh  bbvh-cuda/lbvh.h:13-34,
                                              bbvh-cuda/bbvh-cuda.h:29-73
  RayBuffers        gpu_ray_generator         librta/cuda.h:90-107
  ShirleyRayGenerator cuda::cam_ray_generator_shirley  librta/cuda.h:146-158
  GpuBouncer        gpu_ray_bouncer           librta/cuda.h:193-205
  CisTracer         bbvh_gpu_cis_tracer       bbvh-cuda/bbvh-cuda.h:211-235
                    over gpu_raytracer        librta/cuda.h:279-304
  DiffuseBouncer    a bouncer whose trace_further_bounces() is true once

Every method calls the HIP C-ABI; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .abi import (FLT_MAX, HIT_DTYPE, NODE_DTYPE, RTA_E_INVALID, RTA_E_STACK,
                  RTA_LBVH_FRONTIER_OVERFLOW, TRI_DTYPE, RtaError, RtaTimes, TraceConfig, call,
                  host, ptr)
from .device import DeviceBuffer


class BinaryLbvh:
    """The acceleration structure: triangles in the tree's order, the nodes
    (2n-1 for an LBVH, fewer for a median tree) and the index buffer (tree
    position -> input triangle)."""

    def __init__(self, triangle_data: DeviceBuffer, n: int, node_data: DeviceBuffer,
                 index_buffer: DeviceBuffer, n_nodes: int | None = None,
                 name: str = "lbvh (Karras) on gfx950"):
        self.triangle_data = triangle_data
        self.n = n
        self.node_data = node_data
        self.index_buffer = index_buffer
        self.n_nodes = 2 * n - 1 if n_nodes is None else n_nodes
        self.name = name
        self._packed = None

    def identification(self) -> str:
        return self.name

    def triangle_ptr(self) -> int:
        return self.triangle_data.ptr

    def triangle_count(self) -> int:
        return self.n

    def node_count(self) -> int:
        return self.n_nodes

    def canonical_triangle_ptr(self) -> np.ndarray:
        """Host copy of the AS triangles (bbvh-cuda.h:65-67 downloads too)."""
        return self.triangle_data.download(TRI_DTYPE, self.n)

    def nodes(self) -> np.ndarray:
        return self.node_data.download(NODE_DTYPE, self.node_count())

    def indices(self) -> np.ndarray:
        return self.index_buffer.download(np.uint32, self.n)

    def packed(self, stream=None) -> "PackedBvh":
        if self._packed is None:
            self._packed = PackedBvh.from_nodes(self.node_data, self.node_count(),
                                                self.triangle_data, self.n, stream)
        return self._packed

    def invalidate(self) -> None:
        self._packed = None


class PackedBvh:
    """Child-pair records (64 B per inner node) and vertex-only triangles."""

    def __init__(self, pairs: DeviceBuffer, tri4: DeviceBuffer, flags: DeviceBuffer,
                 root_link: int, root_index: int):
        self.pairs, self.tri4, self.flags = pairs, tri4, flags
        self.root_link, self.root_index = root_link, root_index

    @classmethod
    def from_nodes(cls, node_data: DeviceBuffer, n_nodes: int, tris: DeviceBuffer, n_tris: int,
                   stream=None) -> "PackedBvh":
        pairs = DeviceBuffer(64 * n_nodes)
        tri4 = DeviceBuffer(48 * n_tris)
        flags = DeviceBuffer(4)
        call("rta_bvh_pack", ptr(pairs), ptr(tri4), ptr(flags), ptr(node_data), n_nodes, ptr(tris),
             n_tris, ptr(stream))
        root = node_data.download(NODE_DTYPE, 1)[0]
        return cls(pairs, tri4, flags, int(root["type_left_elems"]), int(root["right_tris"]))

    def exact_division_path(self) -> bool:
        """True when the boxes admit the division-free slab test."""
        return (int(self.flags.download(np.uint32, 1)[0]) & 1) == 0


class LbvhConstructor:
    """Builds a BinaryLbvh from device triangles (expects_host_triangles() is
    false, bbvh-cuda/lbvh.h:213).  Like the reference, the triangle array
    handed in becomes the acceleration structure's triangle array and is
    reordered in place; buffers are reused while n stays the same
    (lbvh.h:157-162)."""

    def __init__(self, verbose: bool = False):
        self.verbose = verbose
        self.last_times: dict | None = None
        self.last_status = 0
        self._n = 0
        self._ws = self._nodes = self._idx = self._status = None

    def identification(self) -> str:
        return "lbvh constructor (bbvh-hip)"

    def expects_host_triangles(self) -> bool:
        return False

    def _setup(self, n: int) -> None:
        if self._ws is not None and self._n == n:
            return
        sz = C.c_size_t(0)
        call("rta_lbvh_workspace_bytes", n, C.byref(sz))
        self._ws = DeviceBuffer(sz.value)
        self._nodes = DeviceBuffer(32 * (2 * n - 1))
        self._idx = DeviceBuffer(4 * n)
        self._status = DeviceBuffer(4)
        self._n = n

    def build(self, tris: DeviceBuffer, n: int, stream=None, timed: bool = True) -> BinaryLbvh:
        self._setup(n)
        times = RtaTimes()
        call("rta_lbvh_build", ptr(tris), n, ptr(self._nodes), ptr(self._idx), ptr(self._ws),
             C.c_size_t(self._ws.nbytes), ptr(self._status),
             C.byref(times) if timed else None, ptr(stream))
        if timed:
            self.last_times = times.as_dict()
            self.last_status = int(self._status.download(np.uint32, 1)[0])
            if self.last_status & RTA_LBVH_FRONTIER_OVERFLOW:
                raise RtaError(RTA_E_INVALID, "rta_lbvh_build",
                               "bottom-up frontier overflow: upper boxes not merged")
            if self.verbose:
                print("lbvh stage times [ms]:", self.last_times)
        return BinaryLbvh(tris, n, self._nodes, self._idx)

    def ranges(self, stream=None) -> DeviceBuffer:
        """Karras ranges [from, to] (sorted positions) of the inner nodes of
        the last build, two words per node, for rta_trace_seed_clusters."""
        out = DeviceBuffer(8 * max(self._n - 1, 1))
        call("rta_lbvh_copy_ranges", ptr(out), ptr(self._ws), self._n, ptr(stream))
        return out


class MedianConstructor:
    """The spatial-median tree built on the host and uploaded, the `-b
    median` builder of the bbvh-cuda plugin (bbvh-cuda/main.cpp:158-163,
    200-235; bbvh/bbvh.h:110-279): expects_host_triangles() is true, the
    structure owns device copies of the nodes and of the triangles in leaf
    order, and the index buffer maps leaf order to input triangles."""

    def identification(self) -> str:
        return "bbvh_constructor_using_median: SM (host, bbvh-hip)"

    def expects_host_triangles(self) -> bool:
        return True

    def build_host(self, tris: np.ndarray):
        """(nodes, triangles in leaf order, input index per leaf position)"""
        n = len(tris)
        if n <= 0:
            raise RtaError(RTA_E_INVALID, "rta_host_median_build", "cannot build 0 triangles")
        tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
        nodes = np.zeros(host().rta_host_median_node_bound(n), NODE_DTYPE)
        out = np.zeros(n, TRI_DTYPE)
        order = np.zeros(n, np.uint32)
        k = host().rta_host_median_build(ptr(tris), n, ptr(nodes), ptr(out), ptr(order))
        if k <= 0:
            raise RtaError(RTA_E_INVALID, "rta_host_median_build", "bad arguments")
        return nodes[:k].copy(), out, order

    def build(self, tris: np.ndarray) -> BinaryLbvh:
        nodes, out, order = self.build_host(tris)
        return BinaryLbvh(DeviceBuffer.from_array(out), len(out), DeviceBuffer.from_array(nodes),
                          DeviceBuffer.from_array(order), n_nodes=len(nodes),
                          name="binary bvh (spatial median, host-built) on gfx950")


class RayBuffers:
    """Device ray arrays: origin[3R], direction[3R], max_t[R]."""

    def __init__(self, n: int):
        self.n = n
        self.gpu_origin = DeviceBuffer(12 * n)
        self.gpu_direction = DeviceBuffer(12 * n)
        self.gpu_maxt = DeviceBuffer(4 * n)

    @classmethod
    def from_host(cls, o: np.ndarray, d: np.ndarray, t: np.ndarray) -> "RayBuffers":
        rb = cls(len(t))
        rb.gpu_origin.upload(o.astype(np.float32))
        rb.gpu_direction.upload(d.astype(np.float32))
        rb.gpu_maxt.upload(t.astype(np.float32))
        return rb

    def download(self, n: int | None = None):
        n = self.n if n is None else n
        return (self.gpu_origin.download(np.float32, 3 * n),
                self.gpu_direction.download(np.float32, 3 * n),
                self.gpu_maxt.download(np.float32, n))


class ShirleyRayGenerator(RayBuffers):
    def __init__(self, w: int, h: int):
        super().__init__(w * h)
        self.w, self.h = w, h
        self.aspect = float(np.float32(w) / np.float32(h))
        self.setup((0, 0, 0), (0, 0, -1), (0, 1, 0), 45.0)

    def setup(self, position, direction, up, fovy) -> None:
        self.position, self.direction, self.up, self.fovy = position, direction, up, fovy

    def generate_rays(self, stream=None) -> None:
        f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
        call("rta_setup_shirley_rays", ptr(self.gpu_direction), ptr(self.gpu_origin),
             ptr(self.gpu_maxt), self.fovy, self.aspect, self.w, self.h, f3(self.direction),
             f3(self.position), f3(self.up), float(FLT_MAX), ptr(stream))


class GpuBouncer:
    """Owns the device intersection buffer (gpu_last_intersection)."""

    def __init__(self, n: int):
        self.n = n
        self.gpu_last_intersection = DeviceBuffer(16 * n)

    def bounce(self) -> None:
        pass

    def trace_further_bounces(self) -> bool:
        return False

    def new_pass(self) -> None:
        pass

    def intersections(self, n: int | None = None) -> np.ndarray:
        return self.gpu_last_intersection.download(HIT_DTYPE, self.n if n is None else n)


class CisTracer:
    """Closest-hit child-is tracer.  layout 'plain' traverses the 32-byte
    node array directly (the reference kernel's data), 'packed' the derived
    child-pair layout, 'persistent' the packed layout with machine-filling
    waves that refill idle lanes; all give identical results."""

    LAYOUTS = ("plain", "packed", "persistent")
    PACKED_ENTRY = "rta_trace_cis_packed"
    PERSISTENT_ENTRY = "rta_trace_cis_persistent"

    def __init__(self, raygen: RayBuffers, bvh: BinaryLbvh, bouncer: GpuBouncer,
                 layout: str = "persistent", stream=None, config: TraceConfig | None = None):
        if layout not in self.LAYOUTS:
            raise ValueError(f"layout must be one of {self.LAYOUTS}")
        self.raygen, self.bvh, self.bouncer, self.layout = raygen, bvh, bouncer, layout
        self.stream = stream
        self.config = config
        self.timings: list[float] = []
        self._err = DeviceBuffer(4)
        self._err.fill_u32(0)
        self._queue = DeviceBuffer(4)

    def identification(self) -> str:
        return f"bbvh-hip cis tracer ({self.layout})"

    def supports_max_t(self) -> bool:
        return True

    def acceleration_structure(self, bvh: BinaryLbvh) -> None:
        self.bvh = bvh

    def prepare_trace(self) -> None:
        call("rta_reset_intersections", ptr(self.bouncer.gpu_last_intersection), self.raygen.n, 1,
             ptr(self.stream))

    def launch(self, n_rays: int | None = None, stats: DeviceBuffer | None = None,
               ray_count: DeviceBuffer | None = None) -> None:
        n = self.raygen.n if n_rays is None else n_rays
        rg, hits = self.raygen, self.bouncer.gpu_last_intersection
        if self.layout == "plain" or stats is not None:
            call("rta_trace_cis", ptr(self.bvh.triangle_data), self.bvh.n, ptr(self.bvh.node_data),
                 ptr(rg.gpu_origin), ptr(rg.gpu_direction), ptr(rg.gpu_maxt), n, 1, ptr(hits),
                 ptr(stats), ptr(self._err), ptr(self.stream))
        elif self.layout == "packed":
            pk = self.bvh.packed(self.stream)
            call(self.PACKED_ENTRY, ptr(pk.tri4), ptr(pk.pairs), ptr(pk.flags), pk.root_link,
                 pk.root_index,
                 ptr(rg.gpu_origin), ptr(rg.gpu_direction), ptr(rg.gpu_maxt), n, ptr(ray_count),
                 ptr(hits), ptr(self._err), ptr(self.stream))
        else:
            pk = self.bvh.packed(self.stream)
            cfg = C.byref(self.config) if self.config is not None else None
            call(self.PERSISTENT_ENTRY, ptr(pk.tri4), ptr(pk.pairs), ptr(pk.flags),
                 pk.root_link, pk.root_index, ptr(rg.gpu_origin), ptr(rg.gpu_direction),
                 ptr(rg.gpu_maxt), n, ptr(ray_count), ptr(hits), ptr(self._err), ptr(self._queue),
                 cfg, ptr(self.stream))

    def check_errors(self) -> None:
        err = int(self._err.download(np.uint32, 1)[0])
        if err:
            self._err.fill_u32(0)
            raise RtaError(RTA_E_STACK, "trace", f"{err} rays overflowed the traversal stack")

    def trace_rays(self) -> float:
        """Wall time of launch + kernel + synchronisation in ms, like the
        reference's wall_time_timer around trace_cis (bbvh-cuda.h:221-229)."""
        t0 = time.perf_counter()
        self.launch()
        call("rta_hip_device_synchronize")
        ms = (time.perf_counter() - t0) * 1e3
        self.check_errors()
        return ms

    def trace(self) -> None:
        """basic_raytracer::trace, librta/raytrav.h:411-421."""
        if hasattr(self.raygen, "generate_rays"):
            self.raygen.generate_rays(self.stream)
        self.bouncer.new_pass()
        self.timings = []
        while True:
            self.prepare_trace()
            call("rta_hip_device_synchronize")
            self.timings.append(self.trace_rays())
            self.bouncer.bounce()
            if not self.bouncer.trace_further_bounces():
                break


class CisShadowTracer(CisTracer):
    """bbvh_gpu_cis_shadow_tracer (bbvh-cuda/bbvh-cuda.h:265-291): the
    closest-hit traversal that ends at the first triangle a leaf accepts
    (trace_shadow_cis, kernels.cu:857-939).  Packed and persistent layouts."""

    LAYOUTS = ("packed", "persistent")
    PACKED_ENTRY = "rta_trace_shadow_cis_packed"
    PERSISTENT_ENTRY = "rta_trace_shadow_cis_persistent"

    def identification(self) -> str:
        return f"bbvh-hip cis shadow tracer ({self.layout})"

    def launch(self, n_rays: int | None = None, stats: DeviceBuffer | None = None,
               ray_count: DeviceBuffer | None = None) -> None:
        if stats is not None:
            raise ValueError("the shadow tracer has no instrumented form")
        super().launch(n_rays, None, ray_count)


def shadow_rays(hits: DeviceBuffer, tris: DeviceBuffer, rays: RayBuffers, n_rays: int, light,
                out: RayBuffers | None = None, stream=None):
    """One shadow ray per hit towards a point light (see rta_shadow_rays).
    Returns the ray buffers, the source ray index of each and the count."""
    out = RayBuffers(n_rays) if out is None else out
    src = DeviceBuffer(4 * n_rays)
    cnt = DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_bounce_work_bytes", n_rays, C.byref(sz))
    work = DeviceBuffer(sz.value)
    lp = (C.c_float * 3)(*[float(x) for x in light])
    call("rta_shadow_rays", ptr(out.gpu_origin), ptr(out.gpu_direction), ptr(out.gpu_maxt),
         ptr(src), ptr(cnt), ptr(hits), ptr(tris), ptr(rays.gpu_origin), ptr(rays.gpu_direction),
         n_rays, lp, ptr(work), C.c_size_t(sz.value), ptr(stream))
    count = int(cnt.download(np.uint32, 1)[0])
    return out, src, count


def diffuse_bounce(hits: DeviceBuffer, tris: DeviceBuffer, rays: RayBuffers, n_rays: int,
                   seed: int, out: RayBuffers | None = None, stream=None):
    """One diffuse bounce per hit ray (see rta_diffuse_bounce).  Returns the
    bounce ray buffers, the source ray index of each bounce ray and the
    count."""
    out = RayBuffers(n_rays) if out is None else out
    src = DeviceBuffer(4 * n_rays)
    cnt = DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_bounce_work_bytes", n_rays, C.byref(sz))
    work = DeviceBuffer(sz.value)
    call("rta_diffuse_bounce", ptr(out.gpu_origin), ptr(out.gpu_direction), ptr(out.gpu_maxt),
         ptr(src), ptr(cnt), ptr(hits), ptr(tris), ptr(rays.gpu_origin), ptr(rays.gpu_direction),
         n_rays, C.c_uint64(seed), ptr(work), C.c_size_t(sz.value), ptr(stream))
    count = int(cnt.download(np.uint32, 1)[0])
    return out, src, count
