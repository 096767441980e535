"""The learned hit predictor (rta_pred_*): the MFMA forward and backward
passes against a float32 torch model of the same network on the CPU (bf16
rounding at the same points, straight-through in the backward pass), and,
trained on bounce rays of a heightfield scene, what it gains over the
untrained network -- with the seeded traversal's records equal to the plain
traversal's on every held-out ray."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import rta_amd
from rta_amd import predictor as P
from rta_amd import scenes
from rta_amd.abi import call, ptr
from rta_amd.device import DeviceBuffer

F32MAX = np.float32(np.finfo(np.float32).max)


def say(msg):
    print(f"[predictor] {msg}", flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def test_param_count_and_layout():
    assert P.param_count() == P.PARAM_COUNT == 83_152
    used = 102 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 * 1025 + 1025
    assert used == 81_537


# ------------------------------------------------------- torch reference

def _torch():
    return pytest.importorskip("torch")


def encode_ref(o, d, lo, hi):
    """The 112-wide bf16 input rows, as float32 (float32 arithmetic for the
    normalisation, float64 sin/cos rounded to float32)."""
    o = o.reshape(-1, 3).astype(np.float32)
    d = d.reshape(-1, 3).astype(np.float32)
    scale = (np.float32(2.0) / (hi - lo)).astype(np.float32)
    v = np.concatenate([((o - lo) * scale - np.float32(1.0)).astype(np.float32), d], axis=1)
    x = np.zeros((len(o), 112), np.float32)
    x[:, :6] = v
    for j in range(6):
        for f in range(8):
            a = np.pi * (np.float32(2.0 ** f) * v[:, j]).astype(np.float64)
            x[:, 6 + 16 * j + 2 * f] = np.sin(a).astype(np.float32)
            x[:, 6 + 16 * j + 2 * f + 1] = np.cos(a).astype(np.float32)
    return x


def model_ref(torch, params, x, label=None, target=None):
    """Forward (and with labels, the loss) of the network in float32 with
    bf16 rounding of inputs, weights and hidden activations."""
    bf = lambda t: t.to(torch.bfloat16).to(torch.float32)
    st = lambda t: t + (bf(t) - t).detach()
    p = {k: torch.tensor(np.array(v), dtype=torch.float32, requires_grad=True)
         for k, v in P.unpack(params).items()}
    h = bf(torch.tensor(x))
    for k in (1, 2, 3):
        h = st(torch.relu(h @ st(p[f"W{k}"]).T + p[f"b{k}"]))
    out = h @ st(p["W4"]).T + p["b4"]
    if label is None:
        return out, p, None
    lab = torch.tensor(label, dtype=torch.int64)
    tgt = torch.tensor(target, dtype=torch.float32)
    mask = lab >= 0
    ce = torch.nn.functional.cross_entropy(out[mask, :1024], lab[mask], reduction="sum")
    se = ((out[mask, 1024] - tgt[mask]) ** 2).sum()
    loss = (ce + se) / len(label)
    loss.backward()
    return out, p, loss


def _random_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.5, 1.5, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o.reshape(-1), d.reshape(-1)


LO = np.float32([-0.5, -0.5, -0.5])
HI = np.float32([1.5, 1.5, 1.5])


@pytest.mark.gpu
def test_forward_matches_reference(gpu):
    torch = _torch()
    n = 517                                    # not a multiple of 16
    o, d = _random_rays(n, 3)
    pred = P.HitPredictor(1024, seed=7, scene_min=LO, scene_max=HI)
    params = pred.params()
    # non-zero biases exercise the bias path
    rng = np.random.default_rng(4)
    for name in ("b1", "b2", "b3", "b4"):
        off, rows, _ = P.LAYOUT[name]
        params[off:off + rows] = rng.normal(scale=0.1, size=rows).astype(np.float32)
    params[P.LAYOUT["b4"][0] + 1025:P.LAYOUT["b4"][0] + 1040] = 0.0
    pred.set_params(params)
    ob, db = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d)
    out = DeviceBuffer(4 * 1040 * n)
    top = DeviceBuffer(16 * n)
    pred.forward(ob, db, n, out=out, top=top)
    rta_amd.synchronize()
    got = out.download(np.float32, 1040 * n).reshape(n, 1040)
    ref, _, _ = model_ref(torch, params, encode_ref(o, d, LO, HI))
    ref = ref.detach().numpy()
    err = np.linalg.norm(got[:, :1025] - ref[:, :1025]) / np.linalg.norm(ref[:, :1025])
    say(f"forward relative L2 error {err:.2e}")
    assert err < 1e-2                           # bf16 activations, fp32 accumulation
    assert np.all(got[:, 1025:] == 0.0)
    # top-4: the GPU's choice among its own logits
    t4 = top.download(np.int32, 4 * n).reshape(n, 4)
    exp = np.argsort(-got[:, :1024], axis=1, kind="stable")[:, :4]
    assert np.array_equal(t4, exp)


@pytest.mark.gpu
def test_gradients_match_reference(gpu):
    torch = _torch()
    n = 700
    o, d = _random_rays(n, 5)
    rng = np.random.default_rng(6)
    label = rng.integers(0, 1024, size=n).astype(np.int32)
    label[rng.uniform(size=n) < 0.3] = -1
    target = rng.normal(size=n).astype(np.float32)
    pred = P.HitPredictor(1024, seed=11, scene_min=LO, scene_max=HI)
    params = pred.params()
    ob, db = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d)
    lb, tb = DeviceBuffer.from_array(label), DeviceBuffer.from_array(target)
    loss = DeviceBuffer(4)
    loss.fill_u32(0)
    call("rta_pred_zero_grads", pred.h, None)
    pred.forward(ob, db, n)
    call("rta_pred_backward", pred.h, ptr(lb), ptr(tb), n, ptr(loss), None)
    rta_amd.synchronize()
    g = P.unpack(pred.grads())
    _, p, lref = model_ref(torch, params, encode_ref(o, d, LO, HI), label, target)
    gl = float(loss.download(np.float32, 1)[0])
    say(f"loss gpu {gl:.6f} ref {float(lref):.6f}")
    assert abs(gl - float(lref)) < 1e-2 * abs(float(lref))
    for name in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"):
        ref = p[name].grad.numpy()
        err = np.linalg.norm(g[name] - ref) / np.linalg.norm(ref)
        say(f"grad {name}: relative L2 error {err:.2e}")
        assert err < 5e-2, name                 # bf16 gradients through the layers
    # padding stays without gradient
    assert np.all(g["W1"][:, 102:] == 0.0) and np.all(g["W4"][1025:] == 0.0)


@pytest.mark.gpu
def test_training_gain_and_exact_seeded_traversal(gpu):
    """Trained on the diffuse bounce rays of four seeds of a 93k-triangle
    heightfield scene and evaluated on a fifth: the top-4 cluster proposals
    hold the closest hit far more often than the untrained network's, and
    the traversal with the seeded max_t gives the plain traversal's records
    on every held-out ray."""
    tris_h = scenes.terrain_scene(grid=192, n_random=20_000)
    n = len(tris_h)
    allv = np.concatenate([tris_h["a"], tris_h["b"], tris_h["c"]])
    lo, hi = allv.min(axis=0), allv.max(axis=0)
    ctor = rta_amd.LbvhConstructor()
    bvh = ctor.build(DeviceBuffer.from_array(tris_h), n)
    ranges = ctor.ranges()
    packed = bvh.packed()
    cam = scenes.terrain_camera(640, 360)
    prim = rta_amd.ShirleyRayGenerator(cam.w, cam.h)
    prim.setup(cam.pos, cam.view_dir, cam.up, cam.fovy)
    prim.generate_rays()
    R = cam.w * cam.h
    bp = rta_amd.GpuBouncer(R)
    tr = rta_amd.CisTracer(prim, bvh, bp)
    tr.prepare_trace(); tr.launch(); rta_amd.synchronize(); tr.check_errors()

    def bounce_set(seed):
        rays, _, cnt = rta_amd.diffuse_bounce(bp.gpu_last_intersection, bvh.triangle_data, prim, R,
                                              seed)
        o, d, t = rays.download(cnt)
        b = rta_amd.GpuBouncer(cnt)
        tb = rta_amd.CisTracer(rta_amd.RayBuffers.from_host(o, d, t), bvh, b)
        tb.prepare_trace(); tb.launch(); rta_amd.synchronize(); tb.check_errors()
        return o, d, t, b.intersections()

    sets = [bounce_set(s) for s in (101, 102, 103, 104, 105)]
    o_tr = np.concatenate([s[0] for s in sets[:4]])
    d_tr = np.concatenate([s[1] for s in sets[:4]])
    lab_tr, tgt_tr = P.cluster_labels(np.concatenate([s[3] for s in sets[:4]]), n)
    n_tr = len(lab_tr)
    o_ev, d_ev, t_ev, h_ev = sets[4]
    lab_ev, _ = P.cluster_labels(h_ev, n)
    n_ev = len(lab_ev)
    say(f"{n} triangles, {n_tr} training rays, {n_ev} held-out rays, "
        f"hit share {np.mean(lab_ev >= 0):.2f}")

    def accuracy(pred):
        top = DeviceBuffer(16 * n_ev)
        pred.forward(DeviceBuffer.from_array(o_ev), DeviceBuffer.from_array(d_ev), n_ev, top=top)
        rta_amd.synchronize()
        return P.top4_accuracy(top.download(np.int32, 4 * n_ev).reshape(-1, 4), lab_ev)

    untrained = P.HitPredictor(max(n_ev, 8192), seed=3, scene_min=lo, scene_max=hi)
    acc0 = accuracy(untrained)
    pred = P.HitPredictor(max(n_ev, 8192), seed=3, scene_min=lo, scene_max=hi)
    losses, secs = pred.train(DeviceBuffer.from_array(o_tr), DeviceBuffer.from_array(d_tr),
                              DeviceBuffer.from_array(lab_tr), DeviceBuffer.from_array(tgt_tr),
                              n_tr, epochs=20, batch=8192, lr=2e-3)
    acc1 = accuracy(pred)
    say(f"losses {['%.3f' % l for l in losses[::5]]} in {secs:.1f} s; top-4 accuracy "
        f"untrained {acc0:.3f} trained {acc1:.3f} (chance {4 / 1024:.4f})")
    assert losses[-1] < 0.9 * losses[0]
    # the trained proposals hold the closest hit tens of times more often
    # than the untrained network's (which are near chance, 4/1024)
    assert acc1 > 0.1 and acc1 > 10 * max(acc0, 4 / 1024)

    # seeded traversal of the held-out rays
    rays = rta_amd.RayBuffers.from_host(o_ev, d_ev, t_ev)
    top, t_seed, maxt2 = DeviceBuffer(16 * n_ev), DeviceBuffer(4 * n_ev), DeviceBuffer(4 * n_ev)
    err = DeviceBuffer(4)
    err.fill_u32(0)
    P.seed_max_t(pred, packed, ranges, n, rays, n_ev, top, t_seed, maxt2, err)
    rta_amd.synchronize()
    assert int(err.download(np.uint32, 1)[0]) == 0
    seeds = t_seed.download(np.float32, n_ev)
    t2 = maxt2.download(np.float32, n_ev)
    hit = h_ev["t"] != F32MAX
    # a seed is a real hit, never before the closest one
    assert np.all(seeds[hit] >= h_ev["t"][hit])
    assert np.all(seeds[~hit] == F32MAX)
    say(f"rays with a seed {np.mean(seeds != F32MAX):.3f}, "
        f"seed equal to the closest hit {np.mean(seeds[hit] == h_ev['t'][hit]):.3f}")
    b2 = rta_amd.GpuBouncer(n_ev)
    tr2 = rta_amd.CisTracer(rta_amd.RayBuffers.from_host(o_ev, d_ev, t2), bvh, b2)
    tr2.prepare_trace(); tr2.launch(); rta_amd.synchronize(); tr2.check_errors()
    got = b2.intersections()
    assert np.array_equal(got["t"] != F32MAX, hit)
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(got[f][hit]), bits(h_ev[f][hit])), f
    assert np.array_equal(got["ref"][hit], h_ev["ref"][hit])
