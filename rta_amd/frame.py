"""The dynamic frame: per-frame vertex jitter, full LBVH rebuild, primary
rays, diffuse bounces and one shadow ray per hit, all enqueued on one
stream through the C-ABI so that the whole frame can be captured into a
graph once and replayed.

Stage order of one frame (every ray count stays on the device):

    jitter -> build -> pack -> primary rays -> trace (closest hit)
    for level 0..bounces:
        [online fine-tune of a hit predictor on this level's hits, when
         one is attached (OnlineFineTune)]
        shadow rays of level's hits -> trace (shadow)
        if level < bounces: diffuse bounce -> trace (closest hit)
    frame counter += 1

The reference's host loop (rta/main.cpp:389-407, librta/raytrav.h:411-421)
rebuilds with -R and traces one bounce pass per trace_rays(); here the
whole frame is one stream of launches with no host round trip.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import FLT_MAX, HIT_DTYPE, NODE_DTYPE, TRI_DTYPE, TraceConfig, call, ptr
from .bvh import RayBuffers
from .device import DeviceBuffer, Event, Stream


class Graph:
    """A captured stream of launches (rta_hip_graph_*)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    def launch(self, stream: Stream) -> None:
        call("rta_hip_graph_launch", self._h, ptr(stream))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                call("rta_hip_graph_destroy", self._h)
            except Exception:
                pass
            self._h = None


class OnlineFineTune:
    """Per-frame fine-tune of a hit predictor (rta_amd.predictor.HitPredictor)
    on the frame's own plain-traversal hits, enqueued inside the frame after
    the trace of ray level `level` (1 = the first bounce):

        Karras ranges -> treelet cut of this frame's tree (the classes)
        repeat `steps` times:
            sample `batch` rays of the level with their hits (a new hashed
            sample per frame and step) -> labels -> inference (the
            predictions before the step, kept for accuracy) -> fused
            training step (rta_pred_train_step: forward, loss, gradients and
            Adam with the device step counter, three launches)

    Only kernels and device copies: the stage is captured with the frame."""

    def __init__(self, pred, n_tris: int, batch: int = 8192, steps: int = 4, lr: float = 1e-3,
                 level: int = 1, seed: int = 7):
        if batch > pred.max_batch:
            raise ValueError("batch exceeds the predictor's max_batch")
        self.pred, self.batch, self.steps, self.lr, self.level, self.seed = (
            pred, batch, steps, lr, level, seed)
        self.ranges = DeviceBuffer(8 * max(n_tris - 1, 1))
        self.table = DeviceBuffer(32 * 1024)
        self.count = DeviceBuffer(4)
        self.o, self.d = DeviceBuffer(12 * batch), DeviceBuffer(12 * batch)
        self.hits = DeviceBuffer(16 * batch)
        self.label = [DeviceBuffer(4 * batch) for _ in range(steps)]
        self.target = [DeviceBuffer(4 * batch) for _ in range(steps)]
        self.top = DeviceBuffer(16 * batch)
        self.loss = DeviceBuffer(4)
        self.step = DeviceBuffer(4)
        self.step.fill_u32(0)
        self.frame_buf = DeviceBuffer(4)

    def enqueue(self, frame, stream: Stream, frame_word: DeviceBuffer | None) -> None:
        """frame: the enclosing DynamicFrame; frame_word: its device frame
        counter when captured, None for an eager frame (then frame_buf holds
        the frame number)."""
        s = ptr(stream)
        n = frame.n
        call("rta_lbvh_copy_ranges", ptr(self.ranges), ptr(frame.ws), n, s)
        call("rta_treelet_cut", ptr(frame.pairs), frame._root[0], frame._root[1],
             ptr(self.ranges), n, 1024, ptr(self.table), ptr(self.count), s)
        rays, hits, cnt = frame.rays[self.level], frame.hits[self.level], frame.counts[self.level]
        fw = frame_word if frame_word is not None else self.frame_buf
        call("rta_hip_fill_u32", ptr(self.loss), 0, 1, s)
        for k in range(self.steps):
            call("rta_pred_sample", ptr(rays.gpu_origin), ptr(rays.gpu_direction), ptr(hits),
                 ptr(cnt), self.batch, ptr(fw), C.c_uint64(self.seed * 1000003 + k),
                 ptr(self.o), ptr(self.d), ptr(self.hits), s)
            call("rta_treelet_labels", ptr(self.hits), self.batch, ptr(self.table),
                 ptr(self.count), ptr(self.label[k]), ptr(self.target[k]), s)
            if k == 0:
                self.pred.infer(self.o, self.d, self.batch, top=self.top, stream=stream)
            call("rta_pred_train_step", self.pred.h, ptr(self.o), ptr(self.d),
                 ptr(self.label[k]), ptr(self.target[k]), self.batch, self.pred.scene_min,
                 self.pred.scene_max, C.c_float(self.lr), ptr(self.step), ptr(self.loss), s)

    def accuracy(self) -> float:
        """Top-4 accuracy, on the first sample of the last frame, of the
        predictions made before that frame's steps."""
        from .predictor import top4_accuracy

        top = self.top.download(np.int32, 4 * self.batch).reshape(-1, 4)
        return top4_accuracy(top, self.label[0].download(np.int32, self.batch))

    def mean_loss(self) -> float:
        """Mean of the last frame's per-step losses."""
        return float(self.loss.download(np.float32, 1)[0]) / self.steps


class DynamicFrame:
    """Buffers and launch sequence of one frame of the dynamic scene.
    primary_stack: LDS stack entries of the primary trace (16, 24 or 32;
    rays that need more are traced again with 32 entries, so the records do
    not change; coherent rays can gain from a short stack's extra resident
    waves, the dynamic scene's measured none); the other traces use
    config's stack.  concurrent_shadows: the shadow rays of each level are
    generated and traced on a second stream, beside the next bounce (both
    only read that level's hits), and an attached fine-tune stage runs on a
    third (it reads its level's rays and hits, the pair records and the
    build's ranges, which nothing later in the frame writes); timed frames
    keep every stage in sequence for their per-stage events."""

    def __init__(self, verts: np.ndarray, idx: np.ndarray, camera, light, sigma: float,
                 seed: int, bounces: int = 2, shadows: bool = True,
                 config: TraceConfig | None = None, stream: Stream | None = None,
                 primary_stack: int = 32, concurrent_shadows: bool = True):
        self.n = len(idx) // 3
        if self.n < 2:
            raise ValueError("the dynamic frame needs at least two triangles")
        self.cam, self.light = camera, tuple(float(x) for x in light)
        self.sigma, self.seed, self.bounces, self.shadows = sigma, seed, bounces, shadows
        self.config = config if config is not None else TraceConfig(64, 16, 32, 0, 10)
        self.primary_config = TraceConfig(*self.config.as_tuple()[:5], primary_stack)
        self.stream = stream if stream is not None else Stream()
        n, R = self.n, camera.w * camera.h
        self.R = R
        self.verts = DeviceBuffer.from_array(np.ascontiguousarray(verts, np.float32))
        self.idx = DeviceBuffer.from_array(np.ascontiguousarray(idx, np.uint32))
        self.tris = DeviceBuffer(104 * n)
        self.nodes = DeviceBuffer(32 * (2 * n - 1))
        self.index = DeviceBuffer(4 * n)
        self.status = DeviceBuffer(4)
        sz = C.c_size_t(0)
        call("rta_lbvh_workspace_bytes", n, C.byref(sz))
        self.ws, self.ws_bytes = DeviceBuffer(sz.value), sz.value
        self.pairs = DeviceBuffer(64 * (2 * n - 1))
        self.tri4 = DeviceBuffer(48 * n)
        self.flags = DeviceBuffer(4)
        # level k rays / hits (k = 0 primary), shadow rays / hits per level
        self.rays = [RayBuffers(R) for _ in range(bounces + 1)]
        self.hits = [DeviceBuffer(16 * R) for _ in range(bounces + 1)]
        self.counts = [DeviceBuffer(4) for _ in range(bounces + 1)]
        self.srays = [RayBuffers(R) for _ in range(bounces + 1)] if shadows else []
        self.shits = [DeviceBuffer(16 * R) for _ in range(bounces + 1)] if shadows else []
        self.scounts = [DeviceBuffer(4) for _ in range(bounces + 1)] if shadows else []
        self.src = DeviceBuffer(4 * R)
        call("rta_bounce_work_bytes", R, C.byref(sz))
        self.bwork, self.bwork_bytes = DeviceBuffer(sz.value), sz.value
        self.err = DeviceBuffer(4)
        self.err.fill_u32(0)
        self.queue = DeviceBuffer(4)
        self.overflow = DeviceBuffer(4 * (R + 1))   # short-stack rays traced again
        self.frame_word = DeviceBuffer(4)
        self.frame_word.fill_u32(0)
        # the second stream of the shadow traces, with its own scratch
        self.concurrent = bool(concurrent_shadows and shadows)
        if self.concurrent:
            self.side = Stream()
            self.side_queue = DeviceBuffer(4)
            self.side_overflow = DeviceBuffer(4 * (R + 1))
            self.side_src = DeviceBuffer(4 * R)
            self.side_bwork = DeviceBuffer(self.bwork_bytes)
            self.ev_ready = [Event() for _ in range(bounces + 1)]
            self.ev_side_done = Event()
            self.ft_stream = Stream()
            self.ev_ft_ready, self.ev_ft_done = Event(), Event()
        self.counts[0].fill_u32(R)
        self.events: list[Event] = []
        self._root = None
        self.finetune: OnlineFineTune | None = None
        # the fills above ran on the default stream; the frame's own stream
        # does not wait for it
        call("rta_hip_device_synchronize")

    # ------------------------------------------------------------ stages

    def _trace(self, rays: RayBuffers, count: DeviceBuffer, hits: DeviceBuffer,
               shadow: bool, config: TraceConfig | None = None, side: bool = False) -> None:
        s = ptr(self.side if side else self.stream)
        call("rta_reset_intersections", ptr(hits), self.R, 1, s)
        entry = "rta_trace_shadow_cis_persistent_ovf" if shadow else "rta_trace_cis_persistent_ovf"
        cfg = config if config is not None else self.config
        queue, overflow = ((self.side_queue, self.side_overflow) if side
                           else (self.queue, self.overflow))
        call(entry, ptr(self.tri4), ptr(self.pairs), ptr(self.flags), self._root[0], self._root[1],
             ptr(rays.gpu_origin), ptr(rays.gpu_direction), ptr(rays.gpu_maxt), self.R,
             ptr(count), ptr(hits), ptr(self.err), ptr(queue), C.byref(cfg), ptr(overflow), s)

    def _shadows(self, k: int, side: bool) -> None:
        """Shadow rays of level k from its hits, and their trace."""
        s = ptr(self.side if side else self.stream)
        src, bwork = (self.side_src, self.side_bwork) if side else (self.src, self.bwork)
        call("rta_shadow_rays", ptr(self.srays[k].gpu_origin), ptr(self.srays[k].gpu_direction),
             ptr(self.srays[k].gpu_maxt), ptr(src), ptr(self.scounts[k]), ptr(self.hits[k]),
             ptr(self.tris), ptr(self.rays[k].gpu_origin), ptr(self.rays[k].gpu_direction),
             self.R, (C.c_float * 3)(*self.light), ptr(bwork), C.c_size_t(self.bwork_bytes), s)
        self._trace(self.srays[k], self.scounts[k], self.shits[k], shadow=True, side=side)

    def _mark(self, timed: bool) -> None:
        if timed:
            ev = Event()
            ev.record(self.stream)
            self.events.append(ev)

    def enqueue(self, frame: int | None = None, timed: bool = False) -> None:
        """Enqueue one frame.  frame=None takes the frame number from the
        device counter (and advances it), which is what a captured graph
        does on every replay."""
        if self._root is None:
            raise RuntimeError("call prime() once before enqueueing frames")
        s = ptr(self.stream)
        f3 = lambda v: (C.c_float * 3)(*v)
        self.events = []
        self._mark(timed)
        fw = ptr(self.frame_word) if frame is None else None
        call("rta_jitter_triangles", ptr(self.tris), ptr(self.verts), ptr(self.idx), self.n,
             self.sigma, C.c_uint64(self.seed), 0 if frame is None else frame, fw, s)
        self._mark(timed)
        # the build writes the traversal layout in its own passes
        call("rta_lbvh_build_packed", ptr(self.tris), self.n, ptr(self.nodes), ptr(self.index),
             ptr(self.pairs), ptr(self.tri4), ptr(self.flags), ptr(self.ws),
             C.c_size_t(self.ws_bytes), ptr(self.status), None, s)
        self._mark(timed)
        cam, p = self.cam, self.rays[0]
        call("rta_setup_shirley_rays", ptr(p.gpu_direction), ptr(p.gpu_origin), ptr(p.gpu_maxt),
             cam.fovy, cam.aspect, cam.w, cam.h, f3(cam.view_dir), f3(cam.pos), f3(cam.up),
             float(FLT_MAX), s)
        self._trace(self.rays[0], self.counts[0], self.hits[0], shadow=False,
                    config=self.primary_config)
        self._mark(timed)
        side = self.concurrent and not timed
        ft_forked = False
        for k in range(self.bounces + 1):
            ft = self.finetune
            if ft is not None and k == ft.level:
                if frame is not None:
                    call("rta_hip_fill_u32", ptr(ft.frame_buf), frame, 1, s)
                fw = self.frame_word if frame is None else None
                if side:
                    # level k's rays and hits are ready: the fine-tune on
                    # its own stream, beside the rest of the frame
                    self.ev_ft_ready.record(self.stream)
                    self.ft_stream.wait(self.ev_ft_ready)
                    ft.enqueue(self, self.ft_stream, fw)
                    self.ev_ft_done.record(self.ft_stream)
                    ft_forked = True
                else:
                    ft.enqueue(self, self.stream, fw)
                self._mark(timed)
            if self.shadows:
                if side:
                    # level k's hits are ready: its shadows on the second
                    # stream, beside the next bounce
                    self.ev_ready[k].record(self.stream)
                    self.side.wait(self.ev_ready[k])
                self._shadows(k, side)
                self._mark(timed)
            if k < self.bounces:
                nxt = self.rays[k + 1]
                call("rta_diffuse_bounce", ptr(nxt.gpu_origin), ptr(nxt.gpu_direction),
                     ptr(nxt.gpu_maxt), ptr(self.src), ptr(self.counts[k + 1]),
                     ptr(self.hits[k]), ptr(self.tris), ptr(self.rays[k].gpu_origin),
                     ptr(self.rays[k].gpu_direction), self.R,
                     C.c_uint64(bounce_seed(self.seed, k + 1)), ptr(self.bwork),
                     C.c_size_t(self.bwork_bytes), s)
                self._trace(nxt, self.counts[k + 1], self.hits[k + 1], shadow=False)
                self._mark(timed)
        if side:
            # the frame ends when every stream has
            self.ev_side_done.record(self.side)
            self.stream.wait(self.ev_side_done)
        if ft_forked:
            self.stream.wait(self.ev_ft_done)
        if frame is None:
            call("rta_hip_add_u32", ptr(self.frame_word), 1, s)

    def attach_finetune(self, ft: OnlineFineTune | None) -> None:
        """Enqueue (or, with None, stop enqueueing) an online fine-tune stage
        after the trace of ray level ft.level; graphs captured before keep
        what they captured."""
        if ft is not None and not 0 <= ft.level <= self.bounces:
            raise ValueError("fine-tune level outside the frame's ray levels")
        self.finetune = ft

    def prime(self) -> None:
        """One eager jitter + build to learn the root's link words (an
        inner node for any scene whose Morton codes are not all equal);
        replays bake them in."""
        s = ptr(self.stream)
        call("rta_jitter_triangles", ptr(self.tris), ptr(self.verts), ptr(self.idx), self.n,
             self.sigma, C.c_uint64(self.seed), 0, None, s)
        call("rta_lbvh_build", ptr(self.tris), self.n, ptr(self.nodes), ptr(self.index),
             ptr(self.ws), C.c_size_t(self.ws_bytes), ptr(self.status), None, s)
        self.stream.synchronize()
        r = self.nodes.download(NODE_DTYPE, 1)[0]
        if not (int(r["type_left_elems"]) & 1):
            raise RuntimeError("all Morton codes are equal: the root is a leaf")
        self._root = (int(r["type_left_elems"]), int(r["right_tris"]))

    def capture(self) -> Graph:
        """Capture one frame (frame number from the device counter)."""
        call("rta_hip_graph_begin", ptr(self.stream))
        h = C.c_void_p()
        try:
            self.enqueue(None)
        finally:
            # always end the capture so the stream is usable again
            call("rta_hip_graph_end", ptr(self.stream), C.byref(h))
        return Graph(h)

    # ----------------------------------------------------------- results

    def set_frame_counter(self, frame: int) -> None:
        self.frame_word.fill_u32(frame, 1, self.stream)

    def check_errors(self) -> None:
        e = int(self.err.download(np.uint32, 1)[0])
        if e:
            raise RuntimeError(f"{e} rays overflowed the traversal stack")

    def ray_counts(self) -> dict:
        c = {"primary": self.R}
        for k in range(1, self.bounces + 1):
            c[f"bounce{k}"] = int(self.counts[k].download(np.uint32, 1)[0])
        for k in range(len(self.scounts)):
            c[f"shadow{k}"] = int(self.scounts[k].download(np.uint32, 1)[0])
        return c

    def results(self) -> dict:
        """Host copies of everything a parity check needs."""
        out = {"tris": self.tris.download(TRI_DTYPE, self.n),
               "nodes": self.nodes.download(NODE_DTYPE, 2 * self.n - 1),
               "counts": self.ray_counts()}
        out["hits"] = [h.download(HIT_DTYPE, self.R) for h in self.hits]
        out["rays"] = [r.download() for r in self.rays]
        out["shadow_hits"] = [h.download(HIT_DTYPE, self.R) for h in self.shits]
        out["shadow_rays"] = [r.download() for r in self.srays]
        return out


def bounce_seed(seed: int, level: int) -> int:
    """Seed of the diffuse bounce that produces level `level` rays."""
    return (seed * 0x9E3779B1 + level) & 0xFFFFFFFFFFFFFFFF
