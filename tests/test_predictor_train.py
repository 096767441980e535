"""The fused training step of the hit predictor (rta_pred_train_step and
rta_pred_train_batches: three launches, no atomics) against the float32
torch model of the network (gradients and loss), against the layered path
(rta_pred_forward / backward / adam), and against Adam's closed form at the
first step; its determinism; and that the weight packs it rewrites feed the
fused inference the updated weights."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rta_amd
from rta_amd import predictor as P
from rta_amd.abi import call, ptr
from rta_amd.device import DeviceBuffer

from test_predictor import HI, LO, _random_rays, _torch, encode_ref, model_ref, say

NAMES = ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _batch(n, seed, miss=0.3):
    o, d = _random_rays(n, seed)
    rng = np.random.default_rng(seed + 1)
    label = rng.integers(0, 1024, size=n).astype(np.int32)
    label[rng.uniform(size=n) < miss] = -1
    target = rng.normal(size=n).astype(np.float32)
    return o, d, label, target


def _dev(*arrays):
    return [DeviceBuffer.from_array(a) for a in arrays]


def _biased(pred, seed):
    """Non-zero biases (the padding rows of b4 stay zero)."""
    params = pred.params()
    rng = np.random.default_rng(seed)
    for name in ("b1", "b2", "b3", "b4"):
        off, rows, _ = P.LAYOUT[name]
        params[off:off + rows] = rng.normal(scale=0.1, size=rows).astype(np.float32)
    params[P.LAYOUT["b4"][0] + 1025:P.LAYOUT["b4"][0] + 1040] = 0.0
    pred.set_params(params)
    return params


def _fused_step(pred, bufs, n, lr, step_word, loss):
    ob, db, lb, tb = bufs
    call("rta_pred_train_step", pred.h, ptr(ob), ptr(db), ptr(lb), ptr(tb), n, pred.scene_min,
         pred.scene_max, C.c_float(lr), ptr(step_word), ptr(loss), None)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [700, 4096 + 17])
def test_fused_gradients_match_reference_and_layered_path(gpu, n):
    """One fused step's gradient (lr 0: the parameters stay) against the
    torch model (relative L2 5e-2, as the layered path) and against the
    layered path's gradient of the same batch (1e-2: the same bf16 network
    with other summation orders); the losses agree."""
    torch = _torch()
    o, d, label, target = _batch(n, 5)
    bufs = _dev(o, d, label, target)
    a = P.HitPredictor(8192, seed=11, scene_min=LO, scene_max=HI)
    params = _biased(a, 12)
    step, loss = DeviceBuffer(4), DeviceBuffer(4)
    step.fill_u32(0)
    loss.fill_u32(0)
    _fused_step(a, bufs, n, 0.0, step, loss)
    rta_amd.synchronize()
    assert int(step.download(np.uint32, 1)[0]) == 1
    assert np.array_equal(a.params(), params)            # lr 0
    gf = P.unpack(a.grads())
    lf = float(loss.download(np.float32, 1)[0])

    b = P.HitPredictor(8192, seed=11, scene_min=LO, scene_max=HI)
    b.set_params(params)
    loss_u = DeviceBuffer(4)
    loss_u.fill_u32(0)
    call("rta_pred_zero_grads", b.h, None)
    b.forward(bufs[0], bufs[1], n)
    call("rta_pred_backward", b.h, ptr(bufs[2]), ptr(bufs[3]), n, ptr(loss_u), None)
    rta_amd.synchronize()
    gu = P.unpack(b.grads())
    lu = float(loss_u.download(np.float32, 1)[0])

    _, p, lref = model_ref(torch, params, encode_ref(o, d, LO, HI), label, target)
    lr_ = float(lref.detach())
    say(f"n={n}: loss fused {lf:.6f} layered {lu:.6f} torch {lr_:.6f}")
    assert abs(lf - lr_) < 1e-2 * abs(lr_)
    assert abs(lf - lu) < 1e-4 * abs(lu)
    for name in NAMES:
        ref = p[name].grad.numpy()
        e_ref, e_lay = rel(gf[name], ref), rel(gf[name], gu[name])
        say(f"grad {name}: fused vs torch {e_ref:.2e}, vs layered {e_lay:.2e} "
            f"(layered vs torch {rel(gu[name], ref):.2e})")
        assert e_ref < 5e-2, name
        assert e_lay < 1e-2, name
    assert np.all(gf["W1"][:, 102:] == 0.0) and np.all(gf["W4"][1025:] == 0.0)
    assert np.all(gf["b4"][1025:] == 0.0)
    a.close()
    b.close()


@pytest.mark.gpu
def test_fused_adam_first_step_closed_form(gpu):
    """At step 1 Adam's bias-corrected moments are g and g^2, so every
    parameter moves by -lr g / (|g| + 1e-8): within 5e-5 of lr, which
    covers the float32 rounding of the parameters (|p| < 0.25: 1.5e-8) and
    of the bias corrections 1 / (1 - 0.9) and 1 / (1 - 0.999) (1.3e-5
    relative)."""
    n = 1000
    bufs = _dev(*_batch(n, 21))
    pred = P.HitPredictor(1024, seed=3, scene_min=LO, scene_max=HI)
    p0 = pred.params()
    step, loss = DeviceBuffer(4), DeviceBuffer(4)
    step.fill_u32(0)
    loss.fill_u32(0)
    lr = 1e-3
    _fused_step(pred, bufs, n, lr, step, loss)
    rta_amd.synchronize()
    g = pred.grads()
    p1 = pred.params()
    g64 = g.astype(np.float64)
    exp = p0.astype(np.float64) - lr * g64 / (np.abs(g64) + 1e-8)
    assert np.max(np.abs(p1 - exp)) < 5e-5 * lr
    assert np.count_nonzero(g) > 70_000
    pred.close()


@pytest.mark.gpu
def test_fused_training_is_deterministic_and_batches_equal_steps(gpu):
    """Two predictors from one seed trained by rta_pred_train_batches end
    bit-identical; a third, stepped batch by batch with rta_pred_train_step
    over the same starts (the last batch ragged), too; the loss words
    agree."""
    n, batch = 3 * 2048 + 333, 2048
    o, d, label, target = _batch(n, 31)
    bufs = _dev(o, d, label, target)
    starts = np.array([2048, 6144, 0, 4096], np.int32)
    out = []
    for mode in ("batches", "batches", "steps"):
        pred = P.HitPredictor(batch, seed=5, scene_min=LO, scene_max=HI)
        step, loss = DeviceBuffer(4), DeviceBuffer(4)
        step.fill_u32(0)
        loss.fill_u32(0)
        if mode == "batches":
            call("rta_pred_train_batches", pred.h, ptr(bufs[0]), ptr(bufs[1]), ptr(bufs[2]),
                 ptr(bufs[3]), n, ptr(starts), len(starts), batch, pred.scene_min,
                 pred.scene_max, C.c_float(2e-3), ptr(step), ptr(loss), None)
        else:
            for s0 in starts:
                m = int(min(batch, n - s0))
                call("rta_pred_train_step", pred.h, ptr(bufs[0].ptr + 12 * int(s0)),
                     ptr(bufs[1].ptr + 12 * int(s0)), ptr(bufs[2].ptr + 4 * int(s0)),
                     ptr(bufs[3].ptr + 4 * int(s0)), m, pred.scene_min, pred.scene_max,
                     C.c_float(2e-3), ptr(step), ptr(loss), None)
        rta_amd.synchronize()
        assert int(step.download(np.uint32, 1)[0]) == len(starts)
        out.append((pred.params(), loss.download(np.float32, 1)[0]))
        pred.close()
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[2][0]) and out[0][1] == out[2][1]
    say(f"loss over {len(starts)} steps {float(out[0][1]):.5f}")


@pytest.mark.gpu
def test_packs_follow_fused_steps(gpu):
    """Every fused step rewrites the bf16 copy and both weight packs from the
    updated parameters: after three steps, rebuilding the packs from the
    parameters (set_params) changes neither the fused inference's logits
    (forward pack) nor the next step's gradient (both packs), bit for bit;
    and the inference agrees with the layered forward pass (which reads the
    bf16 copy) to float rounding."""
    n = 1001
    o, d, label, target = _batch(n, 41)
    bufs = _dev(o, d, label, target)
    pred = P.HitPredictor(1024, seed=9, scene_min=LO, scene_max=HI)
    p0 = pred.params()
    step, loss = DeviceBuffer(4), DeviceBuffer(4)
    step.fill_u32(0)
    loss.fill_u32(0)
    for _ in range(3):
        _fused_step(pred, bufs, n, 1e-2, step, loss)
    logits, out = DeviceBuffer(4 * 1040 * n), DeviceBuffer(4 * 1040 * n)
    pred.infer(bufs[0], bufs[1], n, logits=logits)
    _fused_step(pred, bufs, n, 0.0, step, loss)
    rta_amd.synchronize()
    got = logits.download(np.float32, 1040 * n)
    g_steps = pred.grads()
    params = pred.params()
    assert not np.array_equal(params, p0)
    pred.set_params(params)                    # packs rebuilt from the parameters
    pred.infer(bufs[0], bufs[1], n, logits=logits)
    _fused_step(pred, bufs, n, 0.0, step, loss)
    pred.forward(bufs[0], bufs[1], n, out=out)
    rta_amd.synchronize()
    assert np.array_equal(logits.download(np.float32, 1040 * n), got)
    assert np.array_equal(pred.grads(), g_steps)
    err = rel(got, out.download(np.float32, 1040 * n))
    say(f"inference vs layered forward after 3 fused steps: {err:.2e}")
    assert err < 1e-3            # same bf16 network, other summation order


def test_train_batches_rejects_bad_arguments():
    """Argument checks run before any device work (no GPU needed)."""
    lib = rta_amd.hip()
    dummy = C.c_void_p(16)
    starts = np.array([0, 5], np.int32)
    lo = (C.c_float * 3)(0, 0, 0)
    hi = (C.c_float * 3)(1, 1, 1)
    rc = lib.rta_pred_train_batches(None, dummy, dummy, dummy, dummy, 10, ptr(starts), 2, 4, lo,
                                    hi, C.c_float(1e-3), dummy, dummy, None)
    assert rc == rta_amd.abi.RTA_E_INVALID
    rc = lib.rta_pred_train_step(None, dummy, dummy, dummy, dummy, 10, lo, hi, C.c_float(1e-3),
                                 dummy, dummy, None)
    assert rc == rta_amd.abi.RTA_E_INVALID
