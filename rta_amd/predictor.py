"""The learned hit predictor (rta_pred_* in include/rta_hip.h) and the
seeded traversal it feeds.  No counterpart in the reference.

A ray's label is the leaf cluster of its closest hit: cluster
c = sorted position * 1024 // n_tris (the LBVH's triangles are in Morton
order, so a cluster is a compact region), and its distance target log(t).
Inference proposes the 4 best clusters per ray; rta_trace_seed_clusters
finds the closest hit among their triangles, whose t becomes the ray's
max_t for the exact traversal.  Every seed is the t of a real hit, so the
exact traversal sees the closest hit inside its range.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .abi import FLT_MAX, HIT_DTYPE, call, ptr
from .device import DeviceBuffer, Stream

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


def cluster_labels(hits: np.ndarray, n_tris: int):
    """(label, target) for hit records: the cluster of the hit triangle's
    sorted position and log(t); label -1 and target 0 for misses."""
    hit = hits["t"] != FLT_MAX
    label = np.full(len(hits), -1, np.int32)
    label[hit] = (hits["ref"][hit].astype(np.int64) * N_CLUSTERS // n_tris).astype(np.int32)
    target = np.zeros(len(hits), np.float32)
    target[hit] = np.log(np.maximum(hits["t"][hit], np.float32(1e-30)))
    return label, target


def _at(buf: DeviceBuffer, offset_bytes: int):
    return C.c_void_p(buf.ptr + offset_bytes)


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

    def train(self, o, d, label, target, n, epochs=2, batch=8192, lr=2e-3, stream=None,
              seed=0):
        """Adam over `epochs` passes of the n rays in batches (a seeded
        permutation of batch starts per epoch); returns the mean loss per
        epoch and the wall time."""
        rng = np.random.default_rng(seed)
        starts = np.arange(0, n, batch)
        sp = ptr(stream) if stream else None
        losses = []
        t0 = time.perf_counter()
        for _ in range(epochs):
            self.loss.fill_u32(0, stream=stream)
            for s0 in rng.permutation(starts):
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


def seed_max_t(pred: HitPredictor, packed, ranges: DeviceBuffer, n_tris: int, rays, n: int,
               top: DeviceBuffer, t_seed: DeviceBuffer, maxt_out: DeviceBuffer,
               err: DeviceBuffer, stream=None) -> None:
    """Inference (top-4 clusters), the restricted traversal of those
    clusters, and max_t' = min(max_t, seed) for the n rays.  `packed` is a
    PackedBvh (pairs, tri4, flags, root words)."""
    sp = ptr(stream) if stream else None
    for s0 in range(0, n, pred.max_batch):
        m = min(pred.max_batch, n - s0)
        call("rta_pred_forward", pred.h, _at(rays.gpu_origin, 12 * s0),
             _at(rays.gpu_direction, 12 * s0), m, pred.scene_min, pred.scene_max, None,
             _at(top, 16 * s0), None, sp)
    call("rta_trace_seed_clusters", ptr(packed.tri4), ptr(packed.pairs), ptr(packed.flags),
         packed.root_link, packed.root_index, ptr(ranges), n_tris, ptr(rays.gpu_origin),
         ptr(rays.gpu_direction), ptr(rays.gpu_maxt), n, ptr(top), ptr(t_seed), ptr(err), sp)
    call("rta_seed_max_t", ptr(maxt_out), ptr(rays.gpu_maxt), ptr(t_seed), n, sp)


def top4_accuracy(top: np.ndarray, label: np.ndarray) -> float:
    """Share of hit rays whose cluster is among their 4 proposals."""
    hit = label >= 0
    if not hit.any():
        return 0.0
    return float((top[hit] == label[hit, None]).any(axis=1).mean())


__all__ = ["HitPredictor", "cluster_labels", "seed_max_t", "top4_accuracy", "unpack",
           "param_count", "PARAM_COUNT", "N_CLUSTERS", "TOP", "HIT_DTYPE"]
