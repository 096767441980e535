"""The learned hit predictor (rta_pred_* in include/rta_hip.h) and the
seeded traversal it feeds.  No counterpart in the reference.

The tree is cut into at most 1024 treelets (subtrees; rta_treelet_cut).  A
ray's label is the treelet holding its closest hit and its distance target
log(t).  Inference proposes the 4 best treelets per ray;
rta_trace_seed_treelets finds the closest hit among their triangles, whose t
s, widened to s + s * 2^-10, becomes the ray's max_t for the exact
traversal.  Every seed is the t of a real hit, so the closest hit lies at
t <= s; the margin covers the boxes around it, whose computed entry
distance can exceed its computed t by rounding (up to about 2^-16.3
relative on planar triangles met at grazing angles), so the exact traversal
still enters them.
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .abi import HIT_DTYPE, TREELET_DTYPE, call, ptr
from .device import DeviceBuffer, Stream

SEED_MARGIN = np.float32(2.0 ** -10)


def seed_bound(t_seed: np.ndarray) -> np.ndarray:
    """The box-pruning bound a seed imposes, s + s * 2^-10 in float32 (the
    kernels' seed_bound); FLT_MAX widens to +inf."""
    s = np.asarray(t_seed, dtype=np.float32)
    with np.errstate(over="ignore"):
        return (s + s * SEED_MARGIN).astype(np.float32)


N_CLUSTERS = 1024
TOP = 4
OUT_WIDTH = 1040
IN_WIDTH = 112
HIDDEN = 64
# packed parameter layout (floats): name -> (offset, rows, cols)
LAYOUT = {}
_off = 0
for _name, _rows, _cols in (("W1", HIDDEN, IN_WIDTH), ("b1", HIDDEN, 1),
                            ("W2", HIDDEN, HIDDEN), ("b2", HIDDEN, 1),
                            ("W3", HIDDEN, HIDDEN), ("b3", HIDDEN, 1),
                            ("W4", OUT_WIDTH, HIDDEN), ("b4", OUT_WIDTH, 1)):
    LAYOUT[_name] = (_off, _rows, _cols)
    _off += _rows * _cols
PARAM_COUNT = _off


def param_count() -> int:
    n = C.c_int()
    call("rta_pred_param_count", C.byref(n))
    return n.value


def unpack(params: np.ndarray) -> dict:
    """Views of the packed parameter array: W (rows = outputs) and b."""
    out = {}
    for name, (off, rows, cols) in LAYOUT.items():
        a = params[off:off + rows * cols]
        out[name] = a.reshape(rows, cols) if cols > 1 else a
    return out


def _at(buf: DeviceBuffer, offset_bytes: int):
    return C.c_void_p(buf.ptr + offset_bytes)


class TreeletCut:
    """A cut of a packed tree into at most max_treelets subtrees
    (rta_treelet_cut): the device table and count, and labels from hit
    records.  `packed` is a PackedBvh, `ranges` the build's Karras ranges
    (LbvhConstructor.ranges())."""

    def __init__(self, packed, ranges: DeviceBuffer, n_tris: int, max_treelets: int = N_CLUSTERS,
                 stream=None):
        self.max_treelets = max_treelets
        self.table = DeviceBuffer(32 * max_treelets)
        self.count = DeviceBuffer(4)
        call("rta_treelet_cut", ptr(packed.pairs), packed.root_link, packed.root_index,
             ptr(ranges), n_tris, max_treelets, ptr(self.table), ptr(self.count),
             ptr(stream) if stream else None)

    def host(self) -> np.ndarray:
        """The table's entries (TREELET_DTYPE), ascending first position."""
        k = int(self.count.download(np.uint32, 1)[0])
        return self.table.download(TREELET_DTYPE, k)

    def labels(self, hits: DeviceBuffer, n: int, label: DeviceBuffer, target: DeviceBuffer,
               stream=None) -> None:
        """label[i] = treelet of hits[i].ref, target[i] = log(t); -1 / 0 for
        misses (rta_treelet_labels)."""
        call("rta_treelet_labels", ptr(hits), n, ptr(self.table), ptr(self.count), ptr(label),
             ptr(target), ptr(stream) if stream else None)

    def labels_host(self, hits: DeviceBuffer, n: int, stream=None):
        """labels() downloaded: (label int32[n], target float32[n])."""
        lab, tgt = DeviceBuffer(4 * max(n, 1)), DeviceBuffer(4 * max(n, 1))
        self.labels(hits, n, lab, tgt, stream)
        if stream:
            stream.synchronize()
        return lab.download(np.int32, n), tgt.download(np.float32, n)


class HitPredictor:
    """The MLP's device state for batches of up to max_batch rays."""

    def __init__(self, max_batch: int, seed: int = 1, scene_min=(0, 0, 0), scene_max=(1, 1, 1)):
        h = C.c_void_p()
        call("rta_pred_create", max_batch, C.c_uint64(seed), C.byref(h))
        self.h = h
        self.max_batch = max_batch
        self.scene_min = (C.c_float * 3)(*[float(v) for v in scene_min])
        self.scene_max = (C.c_float * 3)(*[float(v) for v in scene_max])
        self.steps = 0
        self.loss = DeviceBuffer(4)
        self.step_word = DeviceBuffer(4)      # the fused step's device step counter

    def close(self):
        if self.h:
            call("rta_pred_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        a = np.empty(PARAM_COUNT, np.float32)
        call("rta_pred_get_params", self.h, ptr(a))
        return a

    def set_params(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == PARAM_COUNT
        call("rta_pred_set_params", self.h, ptr(a))

    CHECKPOINT_FORMAT = "rta-hit-predictor-1"

    def save(self, path: str) -> None:
        """Weight checkpoint: the packed fp32 parameters, the scene box the
        encoding uses and the Adam step count, as a numpy .npz (no pickled
        objects).  Adam's moments are not kept: a loaded network infers
        identically and fine-tunes from fresh moments."""
        np.savez(path, format=np.array(self.CHECKPOINT_FORMAT), params=self.params(),
                 scene_min=np.array(self.scene_min[:], np.float32),
                 scene_max=np.array(self.scene_max[:], np.float32),
                 steps=np.array(self.steps, np.int64))

    @classmethod
    def load(cls, path: str, max_batch: int) -> "HitPredictor":
        """A predictor with the weights and scene box of a checkpoint."""
        with np.load(path, allow_pickle=False) as z:
            if str(z["format"]) != cls.CHECKPOINT_FORMAT:
                raise ValueError(f"{path}: not a hit-predictor checkpoint")
            params = np.asarray(z["params"], np.float32)
            if params.size != PARAM_COUNT:
                raise ValueError(f"{path}: {params.size} parameters, expected {PARAM_COUNT}")
            p = cls(max_batch, scene_min=z["scene_min"], scene_max=z["scene_max"])
            p.set_params(params)
            p.steps = int(z["steps"])
        return p

    def grads(self) -> np.ndarray:
        a = np.empty(PARAM_COUNT, np.float32)
        call("rta_pred_get_grads", self.h, ptr(a))
        return a

    def forward(self, o, d, n, out=None, top=None, logdist=None, stream=None, o_off=0):
        """o, d: device ray arrays (o_off: first ray); out/top/logdist optional
        device outputs."""
        call("rta_pred_forward", self.h, _at(o, 12 * o_off), _at(d, 12 * o_off), n,
             self.scene_min, self.scene_max, ptr(out) if out is not None else None,
             ptr(top) if top is not None else None,
             ptr(logdist) if logdist is not None else None, ptr(stream) if stream else None)

    def infer(self, o, d, n, top=None, logdist=None, logits=None, stream=None, o_off=0,
              top_k=TOP):
        """The fused inference kernel (any n): the top_k (4 or 1) treelets by
        key (4 ids per ray; -1 beyond top_k), the log distance and, for
        testing, all n x 1040 outputs."""
        call("rta_pred_infer", self.h, _at(o, 12 * o_off), _at(d, 12 * o_off), n,
             self.scene_min, self.scene_max, ptr(top) if top is not None else None, top_k,
             ptr(logdist) if logdist is not None else None,
             ptr(logits) if logits is not None else None, ptr(stream) if stream else None)

    def train_step(self, o, d, label, target, n, lr=2e-3, stream=None, o_off=0):
        """One fused step (rta_pred_train_step) on rays [o_off, o_off + n);
        the loss is added to self.loss."""
        self.step_word.fill_u32(self.steps, stream=stream)
        call("rta_pred_train_step", self.h, _at(o, 12 * o_off), _at(d, 12 * o_off),
             _at(label, 4 * o_off), _at(target, 4 * o_off), n, self.scene_min, self.scene_max,
             C.c_float(lr), ptr(self.step_word), ptr(self.loss), ptr(stream) if stream else None)
        self.steps += 1

    def train(self, o, d, label, target, n, epochs=2, batch=8192, lr=2e-3, stream=None,
              seed=0, fused=True):
        """Adam over `epochs` passes of the n rays in batches (a seeded
        permutation of batch starts per epoch); returns the mean loss per
        epoch and the wall time.  fused: the three-launch fused step
        (rta_pred_train_batches, one call per epoch); otherwise forward,
        backward and Adam per batch."""
        rng = np.random.default_rng(seed)
        starts = np.arange(0, n, batch)
        sp = ptr(stream) if stream else None
        losses = []
        t0 = time.perf_counter()
        for _ in range(epochs):
            self.loss.fill_u32(0, stream=stream)
            order = rng.permutation(starts)
            if fused:
                self.step_word.fill_u32(self.steps, stream=stream)
                first = np.ascontiguousarray(order, dtype=np.int32)
                call("rta_pred_train_batches", self.h, ptr(o), ptr(d), ptr(label), ptr(target), n,
                     ptr(first), len(first), batch, self.scene_min, self.scene_max,
                     C.c_float(lr), ptr(self.step_word), ptr(self.loss), sp)
                self.steps += len(first)
            else:
                for s0 in order:
                    m = int(min(batch, n - s0))
                    call("rta_pred_zero_grads", self.h, sp)
                    self.forward(o, d, m, stream=stream, o_off=int(s0))
                    call("rta_pred_backward", self.h, _at(label, 4 * int(s0)),
                         _at(target, 4 * int(s0)), m, ptr(self.loss), sp)
                    self.steps += 1
                    call("rta_pred_adam", self.h, C.c_float(lr), self.steps, sp)
            if stream:
                stream.synchronize()
            losses.append(float(self.loss.download(np.float32, 1)[0]) / len(starts))
        return losses, time.perf_counter() - t0


def seed_max_t(pred: HitPredictor, packed, cut: TreeletCut, rays, n: int, top: DeviceBuffer,
               t_seed: DeviceBuffer, maxt_out: DeviceBuffer, err: DeviceBuffer,
               stream=None, candidates: int = TOP) -> None:
    """Inference (fused kernel; top-1 when one candidate is traversed, else
    top-4), the traversal of the first `candidates` proposed treelets, and
    max_t' = min(max_t, seed + seed * 2^-10) for the n rays.  `packed` is a PackedBvh
    (pairs, tri4, flags)."""
    sp = ptr(stream) if stream else None
    pred.infer(rays.gpu_origin, rays.gpu_direction, n, top=top, stream=stream,
               top_k=1 if candidates == 1 else TOP)
    call("rta_trace_seed_treelets", ptr(packed.tri4), ptr(packed.pairs), ptr(packed.flags),
         ptr(cut.table), ptr(cut.count), ptr(rays.gpu_origin), ptr(rays.gpu_direction),
         ptr(rays.gpu_maxt), n, ptr(top), candidates, ptr(t_seed), ptr(err), sp)
    call("rta_seed_max_t", ptr(maxt_out), ptr(rays.gpu_maxt), ptr(t_seed), n, sp)


def trace_seeded(pred: HitPredictor, packed, cut: TreeletCut, rays, n: int, top: DeviceBuffer,
                 hits: DeviceBuffer, err: DeviceBuffer, queue: DeviceBuffer, config=None,
                 ray_count: DeviceBuffer | None = None, stream=None) -> None:
    """Hit-predicted tracing in two launches: top-1 inference, then the
    persistent traversal that first traverses the proposed treelet and then
    the whole tree under the bound it found
    (rta_trace_cis_persistent_seeded).  `hits` must hold the reset
    sentinel, as for the plain traversal."""
    from .abi import TraceConfig

    cfg = config if config is not None else TraceConfig(64, 16, 32, 0, 10)
    pred.infer(rays.gpu_origin, rays.gpu_direction, n, top=top, stream=stream, top_k=1)
    call("rta_trace_cis_persistent_seeded", ptr(packed.tri4), ptr(packed.pairs),
         ptr(packed.flags), packed.root_link, packed.root_index, ptr(rays.gpu_origin),
         ptr(rays.gpu_direction), ptr(rays.gpu_maxt), n,
         ptr(ray_count) if ray_count is not None else None, ptr(cut.table), ptr(cut.count),
         ptr(top), ptr(hits), ptr(err), ptr(queue), C.byref(cfg),
         ptr(stream) if stream else None)


def top4_accuracy(top: np.ndarray, label: np.ndarray) -> float:
    """Share of hit rays whose treelet is among their 4 proposals."""
    hit = label >= 0
    if not hit.any():
        return 0.0
    return float((top[hit] == label[hit, None]).any(axis=1).mean())


__all__ = ["HitPredictor", "TreeletCut", "seed_bound", "seed_max_t", "trace_seeded",
           "top4_accuracy", "unpack",
           "param_count", "PARAM_COUNT", "N_CLUSTERS", "TOP", "HIT_DTYPE", "TREELET_DTYPE"]
