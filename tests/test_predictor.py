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


def keyed_top4(logits):
    """The fused kernel's selection rule on the CPU: each cluster's key is its
    logit with the low 10 mantissa bits replaced by 1023 - cluster; the 4
    largest keys, largest first."""
    bits = np.ascontiguousarray(logits[:, :1024], np.float32).view(np.uint32)
    keys = ((bits & np.uint32(0xFFFFFC00)) | (np.uint32(1023) - np.arange(1024, dtype=np.uint32)))
    return np.argsort(-keys.view(np.float32), axis=1, kind="stable")[:, :4].astype(np.int32)


def test_keyed_top4_rule():
    """The key rule on hand-made rows: the largest logits win, and logits
    equal in their top 13 mantissa bits go to the lower cluster."""
    rows = np.full((3, 1040), -5.0, np.float32)
    rows[0, [7, 300, 2, 1023]] = [4.0, 3.0, 2.0, 1.0]
    rows[1, [10, 20, 30, 40, 50]] = 1.0              # five equal logits
    base = np.float32(1.5)
    rows[2, 900] = base
    rows[2, 100] = np.nextafter(base, np.float32(2))  # same top bits as 1.5
    rows[2, 5] = np.float32(1.25)
    rows[2, 6] = np.float32(-0.5)
    got = keyed_top4(rows)
    assert got[0].tolist() == [7, 300, 2, 1023]
    assert got[1].tolist() == [10, 20, 30, 40]
    assert got[2].tolist() == [100, 900, 5, 6]


@pytest.mark.gpu
def test_infer_matches_reference(gpu):
    """The fused inference kernel: its outputs against the torch model and the
    unfused forward pass, its top-4 against the key rule on its own logits,
    with a ragged last tile (n not a multiple of 32)."""
    torch = _torch()
    n = 1001
    o, d = _random_rays(n, 13)
    pred = P.HitPredictor(1024, seed=17, scene_min=LO, scene_max=HI)
    params = pred.params()
    rng = np.random.default_rng(14)
    for name in ("b1", "b2", "b3", "b4"):
        off, rows, _ = P.LAYOUT[name]
        params[off:off + rows] = rng.normal(scale=0.1, size=rows).astype(np.float32)
    params[P.LAYOUT["b4"][0] + 1025:P.LAYOUT["b4"][0] + 1040] = 0.0
    pred.set_params(params)
    ob, db = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d)
    logits, top, ld = DeviceBuffer(4 * 1040 * n), DeviceBuffer(16 * n), DeviceBuffer(4 * n)
    pred.infer(ob, db, n, top=top, logdist=ld, logits=logits)
    out = DeviceBuffer(4 * 1040 * n)
    pred.forward(ob, db, n, out=out)
    top2, ld2 = DeviceBuffer(16 * n), DeviceBuffer(4 * n)
    pred.infer(ob, db, n, top=top2, logdist=ld2)
    rta_amd.synchronize()
    got = logits.download(np.float32, 1040 * n).reshape(n, 1040)
    unfused = out.download(np.float32, 1040 * n).reshape(n, 1040)
    ref, _, _ = model_ref(torch, params, encode_ref(o, d, LO, HI))
    ref = ref.detach().numpy()
    err = np.linalg.norm(got[:, :1025] - ref[:, :1025]) / np.linalg.norm(ref[:, :1025])
    err_u = np.linalg.norm(got - unfused) / np.linalg.norm(unfused)
    say(f"fused inference: relative L2 error {err:.2e} vs the torch model, {err_u:.2e} vs the "
        f"unfused forward")
    assert err < 1e-2
    assert err_u < 1e-4                       # same bf16 network, other summation order
    assert np.all(got[:, 1025:] == 0.0)
    t4 = top.download(np.int32, 4 * n).reshape(n, 4)
    assert np.array_equal(t4, keyed_top4(got))
    assert np.array_equal(ld.download(np.float32, n), got[:, 1024])
    # the instance without the logits output gives the same results
    assert np.array_equal(top2.download(np.int32, 4 * n).reshape(n, 4), t4)
    assert np.array_equal(ld2.download(np.float32, n), got[:, 1024])
    # top-1: the best key, the other three entries -1
    top1 = DeviceBuffer(16 * n)
    pred.infer(ob, db, n, top=top1, top_k=1)
    rta_amd.synchronize()
    t1 = top1.download(np.int32, 4 * n).reshape(n, 4)
    assert np.array_equal(t1[:, 0], t4[:, 0])
    assert np.all(t1[:, 1:] == -1)


def _sample_index(seed, frame, i, count):
    """rta_pred_sample's draw on the host (splitmix64 finaliser of
    seed * golden + (frame << 32) + i, modulo count)."""
    m = (1 << 64) - 1
    z = (seed * 0x9E3779B97F4A7C15 + (frame << 32) + i) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return z % count


@pytest.mark.gpu
def test_sample_draws_the_hashed_rays(gpu):
    """rta_pred_sample copies ray i' = hash(seed, frame, i) mod count (origin,
    direction and record) for every output, reads the frame from device
    memory, and gives miss records for an empty set."""
    rng = np.random.default_rng(31)
    n, m, seed = 5000, 777, 12345
    o = rng.normal(size=3 * n).astype(np.float32)
    d = rng.normal(size=3 * n).astype(np.float32)
    hits = np.zeros(n, P.HIT_DTYPE)
    hits["t"] = rng.uniform(size=n).astype(np.float32)
    hits["ref"] = np.arange(n)
    ob, db, hb = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d), DeviceBuffer.from_array(hits)
    cnt, fw = DeviceBuffer(4), DeviceBuffer(4)
    oo, do, ho = DeviceBuffer(12 * m), DeviceBuffer(12 * m), DeviceBuffer(16 * m)
    for count, frame in ((n, 0), (3001, 9), (1, 4), (0, 2)):
        cnt.upload(np.uint32([count]))
        fw.upload(np.uint32([frame]))
        call("rta_pred_sample", ptr(ob), ptr(db), ptr(hb), ptr(cnt), m, ptr(fw), C.c_uint64(seed),
             ptr(oo), ptr(do), ptr(ho), None)
        rta_amd.synchronize()
        got_h = ho.download(P.HIT_DTYPE, m)
        if count == 0:
            assert np.all(got_h["t"] == F32MAX)
            continue
        idx = np.array([_sample_index(seed, frame, i, count) for i in range(m)])
        assert np.array_equal(got_h["ref"], idx)
        assert np.array_equal(oo.download(np.float32, 3 * m).reshape(-1, 3), o.reshape(-1, 3)[idx])
        assert np.array_equal(do.download(np.float32, 3 * m).reshape(-1, 3), d.reshape(-1, 3)[idx])


@pytest.mark.gpu
def test_checkpoint_round_trip(gpu, tmp_path):
    """A weight checkpoint (after a few training steps) loads into a new
    predictor with the same parameters, scene box and step count, which
    infers bit for bit like the original; a foreign file is refused."""
    n = 600
    o, d = _random_rays(n, 51)
    rng = np.random.default_rng(52)
    label = rng.integers(0, 1024, size=n).astype(np.int32)
    target = rng.normal(size=n).astype(np.float32)
    ob, db = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d)
    a = P.HitPredictor(1024, seed=9, scene_min=LO, scene_max=HI)
    a.train(ob, db, DeviceBuffer.from_array(label), DeviceBuffer.from_array(target), n,
            epochs=3, batch=256)
    path = str(tmp_path / "pred.npz")
    a.save(path)
    b = P.HitPredictor.load(path, max_batch=512)
    assert np.array_equal(a.params(), b.params())
    assert b.steps == a.steps == 9
    assert list(b.scene_min) == list(a.scene_min) and list(b.scene_max) == list(a.scene_max)
    outs = []
    for p in (a, b):
        top, ld = DeviceBuffer(16 * n), DeviceBuffer(4 * n)
        p.infer(ob, db, n, top=top, logdist=ld)
        rta_amd.synchronize()
        outs.append((top.download(np.int32, 4 * n), ld.download(np.float32, n)))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(bits(outs[0][1]),
                                                                     bits(outs[1][1]))
    bad = str(tmp_path / "other.npz")
    np.savez(bad, format=np.array("something else"), params=np.zeros(3, np.float32))
    with pytest.raises(ValueError):
        P.HitPredictor.load(bad, max_batch=16)


@pytest.mark.gpu
def test_adam_device_step_matches_host_step(gpu):
    """Adam with the step count in device memory follows Adam with host step
    counts 1, 2, 3 on the same gradients (bias corrections from device powf
    against host powf: relative 1e-5), and advances its counter.  256 rays
    make one accumulation chunk, so both get bit-identical gradients."""
    n = 256
    o, d = _random_rays(n, 41)
    rng = np.random.default_rng(42)
    label = rng.integers(0, 1024, size=n).astype(np.int32)
    target = rng.normal(size=n).astype(np.float32)
    ob, db = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d)
    lb, tb = DeviceBuffer.from_array(label), DeviceBuffer.from_array(target)
    loss, step = DeviceBuffer(4), DeviceBuffer(4)
    step.fill_u32(0)
    a = P.HitPredictor(1024, seed=3, scene_min=LO, scene_max=HI)
    b = P.HitPredictor(1024, seed=3, scene_min=LO, scene_max=HI)
    initial = a.params()
    for k in (1, 2, 3):
        for p, dev in ((a, False), (b, True)):
            call("rta_pred_zero_grads", p.h, None)
            p.forward(ob, db, n)
            call("rta_pred_backward", p.h, ptr(lb), ptr(tb), n, ptr(loss), None)
            if dev:
                call("rta_pred_adam_device", p.h, C.c_float(1e-3), ptr(step), None)
            else:
                call("rta_pred_adam", p.h, C.c_float(1e-3), k, None)
        rta_amd.synchronize()
        pa, pb = a.params(), b.params()
        assert int(step.download(np.uint32, 1)[0]) == k
        assert np.allclose(pa, pb, rtol=1e-5, atol=1e-7), k
        assert not np.array_equal(pa, initial)
        if k == 1:   # same parameters so far, so the same (deterministic) gradients
            assert np.array_equal(a.grads(), b.grads())


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
    say(f"loss gpu {gl:.6f} ref {float(lref.detach()):.6f}")
    assert abs(gl - float(lref.detach())) < 1e-2 * abs(float(lref.detach()))
    for name in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"):
        ref = p[name].grad.numpy()
        err = np.linalg.norm(g[name] - ref) / np.linalg.norm(ref)
        say(f"grad {name}: relative L2 error {err:.2e}")
        assert err < 5e-2, name                 # bf16 gradients through the layers
    # padding stays without gradient
    assert np.all(g["W1"][:, 102:] == 0.0) and np.all(g["W4"][1025:] == 0.0)


# ------------------------------------------------------------- treelets

def cut_restatement(nodes, ranges, n, max_treelets):
    """The treelet cut on the host from the 32-B nodes and Karras ranges:
    rounds that split the inner treelets holding more than half as many
    triangles as the largest inner one (largest first, ties to the lower
    first position), as many as fit; sorted by first position.  Entries are
    (first, word, lo, hi)."""
    big = float(F32MAX)

    def child(c):
        nd = nodes[c]
        if nd["type_left_elems"] & 1:
            a, b = ranges[c]
            return [int(a), (c << 1) | 1, int(b) - int(a) + 1, tuple(nd["min"]), tuple(nd["max"])]
        return [int(nd["right_tris"]), int(nd["right_tris"]) << 1, 1, tuple(nd["min"]),
                tuple(nd["max"])]

    r = nodes[0]
    root_word = 1 if r["type_left_elems"] & 1 else int(r["right_tris"]) << 1
    ent = [[0, root_word, n, (-big,) * 3, (big,) * 3]]
    while len(ent) < max_treelets:
        inner = [e for e in range(len(ent)) if ent[e][1] & 1]
        largest = max((ent[e][2] for e in inner), default=0)
        cand = sorted((e for e in inner if 2 * ent[e][2] > largest),
                      key=lambda e: (-ent[e][2], ent[e][0]))
        splits = min(len(cand), max_treelets - len(ent))
        if splits == 0:
            break
        new = []
        for e in cand[:splits]:
            x = ent[e][1] >> 1
            ent[e] = child(int(nodes[x]["type_left_elems"]) >> 1)
            new.append(child(int(nodes[x]["right_tris"])))
        ent.extend(new)
    return sorted((e[0], e[1], e[3], e[4]) for e in ent)


def _terrain_bvh(grid, n_random):
    tris_h = scenes.terrain_scene(grid=grid, n_random=n_random)
    n = len(tris_h)
    ctor = rta_amd.LbvhConstructor()
    bvh = ctor.build(DeviceBuffer.from_array(tris_h), n)
    return tris_h, n, ctor, bvh


@pytest.mark.gpu
def test_treelet_cut_matches_restatement(gpu):
    """The cut kernel's table equals the host restatement (words, first
    positions and boxes bit for bit) for several treelet limits, and the
    treelets partition the sorted triangle positions."""
    tris_h, n, ctor, bvh = _terrain_bvh(64, 3_000)
    ranges = ctor.ranges()
    packed = bvh.packed()
    rta_amd.synchronize()
    nodes = bvh.node_data.download(rta_amd.abi.NODE_DTYPE, 2 * n - 1)
    rng_h = ranges.download(np.uint32, 2 * (n - 1)).reshape(-1, 2)
    for k in (1, 2, 7, 100, 1024):
        cut = P.TreeletCut(packed, ranges, n, max_treelets=k)
        got = cut.host()
        exp = cut_restatement(nodes, rng_h, n, k)
        say(f"{n} triangles, limit {k}: {len(got)} treelets, largest "
            f"{np.max(np.diff(np.append(got['first'], n)))} triangles")
        assert len(got) == len(exp) and len(got) <= k
        assert got["first"][0] == 0 and np.all(np.diff(got["first"].astype(np.int64)) > 0)
        for g, e in zip(got, exp):
            assert int(g["first"]) == e[0] and int(g["word"]) == e[1]
            assert bits(np.float32(g["lo"])).tobytes() == bits(np.float32(e[2])).tobytes()
            assert bits(np.float32(g["hi"])).tobytes() == bits(np.float32(e[3])).tobytes()
    assert len(got) == 1024                    # 11k triangles allow the full cut


@pytest.mark.gpu
def test_treelet_labels(gpu):
    """Labels from hit records: the treelet holding each ref (a binary search
    on the host) and log t, -1 / 0 for misses, including refs at treelet
    boundaries."""
    tris_h, n, ctor, bvh = _terrain_bvh(48, 1_000)
    cut = P.TreeletCut(bvh.packed(), ctor.ranges(), n)
    table = cut.host()
    firsts = table["first"].astype(np.int64)
    rng = np.random.default_rng(21)
    m = 5000
    hits = np.zeros(m, P.HIT_DTYPE)
    refs = rng.integers(0, n, size=m)
    refs[:len(firsts)] = firsts                      # every boundary
    refs[len(firsts):2 * len(firsts)] = np.maximum(firsts - 1, 0)
    hits["ref"] = refs
    hits["t"] = rng.uniform(1e-3, 10.0, size=m).astype(np.float32)
    miss = rng.uniform(size=m) < 0.2
    hits["t"][miss] = F32MAX
    lab, tgt = cut.labels_host(DeviceBuffer.from_array(hits), m)
    exp = np.searchsorted(firsts, refs, side="right") - 1
    exp[miss] = -1
    assert np.array_equal(lab, exp)
    assert np.all(tgt[miss] == 0.0)
    np.testing.assert_allclose(tgt[~miss], np.log(hits["t"][~miss].astype(np.float64)),
                               rtol=1e-6, atol=1e-6)


@pytest.mark.gpu
def test_training_gain_and_exact_seeded_traversal(gpu):
    """Trained on the diffuse bounce rays of four seeds of a 93k-triangle
    heightfield scene and evaluated on a fifth: the top-4 treelet proposals
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
    cut = P.TreeletCut(packed, ranges, n)
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
    h_tr = np.concatenate([s[3] for s in sets[:4]])
    lab_tr, tgt_tr = cut.labels_host(DeviceBuffer.from_array(h_tr), len(h_tr))
    n_tr = len(lab_tr)
    o_ev, d_ev, t_ev, h_ev = sets[4]
    lab_ev, _ = cut.labels_host(DeviceBuffer.from_array(h_ev), len(h_ev))
    n_ev = len(lab_ev)
    say(f"{n} triangles, {len(cut.host())} treelets, {n_tr} training rays, {n_ev} held-out "
        f"rays, hit share {np.mean(lab_ev >= 0):.2f}")

    def accuracy(pred):
        top = DeviceBuffer(16 * n_ev)
        pred.infer(DeviceBuffer.from_array(o_ev), DeviceBuffer.from_array(d_ev), n_ev, top=top)
        rta_amd.synchronize()
        return P.top4_accuracy(top.download(np.int32, 4 * n_ev).reshape(-1, 4), lab_ev)

    untrained = P.HitPredictor(8192, seed=3, scene_min=lo, scene_max=hi)
    acc0 = accuracy(untrained)
    pred = P.HitPredictor(8192, seed=3, scene_min=lo, scene_max=hi)
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
    P.seed_max_t(pred, packed, cut, rays, n_ev, top, t_seed, maxt2, err)
    rta_amd.synchronize()
    assert int(err.download(np.uint32, 1)[0]) == 0

    # what the trained network gains over the traversal without it (and
    # over the untrained network): inner-node visits per ray, counted by the
    # instrumented plain kernel under each max_t.  A tighter bound never
    # adds a visit, so the comparison holds ray by ray.
    maxt0 = DeviceBuffer(4 * n_ev)
    P.seed_max_t(untrained, packed, cut, rays, n_ev, DeviceBuffer(16 * n_ev),
                 DeviceBuffer(4 * n_ev), maxt0, err)

    def visits(mt):
        stats, scratch = DeviceBuffer(8 * n_ev), DeviceBuffer(16 * n_ev)
        call("rta_reset_intersections", ptr(scratch), n_ev, 1, None)
        call("rta_trace_cis", ptr(bvh.triangle_data), n, ptr(bvh.node_data), ptr(rays.gpu_origin),
             ptr(rays.gpu_direction), ptr(mt), n_ev, 1, ptr(scratch), ptr(stats), ptr(err), None)
        rta_amd.synchronize()
        return stats.download(np.uint32, 2 * n_ev).reshape(-1, 2)[:, 0].astype(np.int64)

    v_plain, v_untrained, v_trained = visits(rays.gpu_maxt), visits(maxt0), visits(maxt2)
    say(f"inner-node visits per ray: plain {v_plain.mean():.2f}, untrained seeds "
        f"{v_untrained.mean():.2f}, trained seeds {v_trained.mean():.2f}")
    assert np.all(v_trained <= v_plain) and np.all(v_untrained <= v_plain)
    assert v_trained.mean() < v_untrained.mean()
    seeds = t_seed.download(np.float32, n_ev)
    t2 = maxt2.download(np.float32, n_ev)
    hit = h_ev["t"] != F32MAX
    # a seed is a real hit, never before the closest one
    assert np.all(seeds[hit] >= h_ev["t"][hit])
    assert np.all(seeds[~hit] == F32MAX)
    # a proposal that holds the closest hit yields it as the seed
    t4 = top.download(np.int32, 4 * n_ev).reshape(-1, 4)
    holds = (t4 == lab_ev[:, None]).any(axis=1) & hit
    assert np.all(seeds[holds] == h_ev["t"][holds])
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

    # the seeding and the exact traversal fused into one persistent launch
    hits_f, queue = DeviceBuffer(16 * n_ev), DeviceBuffer(4)
    call("rta_reset_intersections", ptr(hits_f), n_ev, 1, None)
    P.trace_seeded(pred, packed, cut, rays, n_ev, top, hits_f, err, queue)
    rta_amd.synchronize()
    assert int(err.download(np.uint32, 1)[0]) == 0
    got = hits_f.download(P.HIT_DTYPE, n_ev)
    assert np.array_equal(got["t"] != F32MAX, hit)
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(got[f][hit]), bits(h_ev[f][hit])), f
    assert np.array_equal(got["ref"][hit], h_ev["ref"][hit])
    say("fused seeded persistent traversal: records equal the plain traversal's")


@pytest.mark.gpu
def test_seeded_traversal_exact_on_planar_scene(gpu):
    """Planar triangles met at grazing angles, where a box's computed entry
    distance can exceed the computed t of a triangle inside it: with
    perfect proposals (each ray's own closest-hit treelet, the strongest
    seeds there are) both seeded paths -- the seeding launch followed by the
    traversal under rta_seed_max_t's bound, and the persistent seeded
    traversal -- return the plain records, which equal the oracle's.  The
    oracle test of the same scene shows that a bound at the seed itself
    would lose closest hits there."""
    import oracle
    from helpers import grazing_rays, planar_scene

    tris = planar_scene()
    n = len(tris)
    ob = oracle.lbvh_build(tris)
    o, d, t = grazing_rays(200_000)
    ref, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], o, d, t)
    ctor = rta_amd.LbvhConstructor()
    bvh = ctor.build(DeviceBuffer.from_array(tris), n)
    packed = bvh.packed()
    cut = P.TreeletCut(packed, ctor.ranges(), n)
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    m = len(t)

    def plain(maxt):
        b = rta_amd.GpuBouncer(m)
        tr = rta_amd.CisTracer(rta_amd.RayBuffers.from_host(o, d, maxt), bvh, b)
        tr.prepare_trace(); tr.launch(); rta_amd.synchronize(); tr.check_errors()
        return b.intersections()

    h = plain(t)
    hit = ref["t"] != F32MAX
    say(f"planar: {n} triangles, {m} grazing rays, {int(hit.sum())} hits")
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(h[f][hit]), bits(ref[f][hit])), f
    assert np.array_equal(h["ref"][hit], ref["ref"][hit])
    assert np.all(h["t"][~hit] == F32MAX)

    lab, _ = cut.labels_host(DeviceBuffer.from_array(h), m)
    top = np.full((m, 4), -1, np.int32)
    top[:, 0] = lab                        # -1 (a miss) is never a treelet
    top_d = DeviceBuffer.from_array(top)
    err = DeviceBuffer(4)
    err.fill_u32(0)
    t_seed, maxt2 = DeviceBuffer(4 * m), DeviceBuffer(4 * m)
    call("rta_trace_seed_treelets", ptr(packed.tri4), ptr(packed.pairs), ptr(packed.flags),
         ptr(cut.table), ptr(cut.count), ptr(rays.gpu_origin), ptr(rays.gpu_direction),
         ptr(rays.gpu_maxt), m, ptr(top_d), 1, ptr(t_seed), ptr(err), None)
    call("rta_seed_max_t", ptr(maxt2), ptr(rays.gpu_maxt), ptr(t_seed), m, None)
    rta_amd.synchronize()
    assert int(err.download(np.uint32, 1)[0]) == 0
    seeds = t_seed.download(np.float32, m)
    # the perfect proposal yields the closest hit itself as the seed
    assert np.array_equal(seeds[hit], h["t"][hit]) and np.all(seeds[~hit] == F32MAX)
    mt = maxt2.download(np.float32, m)
    assert np.array_equal(bits(mt), bits(np.minimum(t, P.seed_bound(seeds))))
    h2 = plain(mt)
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(bits(h2[f][hit]), bits(h[f][hit])), f

    hits = DeviceBuffer(16 * m)
    call("rta_reset_intersections", ptr(hits), m, 1, None)
    queue = DeviceBuffer(4)
    call("rta_trace_cis_persistent_seeded", ptr(packed.tri4), ptr(packed.pairs),
         ptr(packed.flags), packed.root_link, packed.root_index, ptr(rays.gpu_origin),
         ptr(rays.gpu_direction), ptr(rays.gpu_maxt), m, None, ptr(cut.table), ptr(cut.count),
         ptr(top_d), ptr(hits), ptr(err), ptr(queue), None, None)
    rta_amd.synchronize()
    assert int(err.download(np.uint32, 1)[0]) == 0
    h3 = hits.download(rta_amd.HIT_DTYPE, m)
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(bits(h3[f][hit]), bits(h[f][hit])), f
    assert np.all(h3["t"][~hit] == F32MAX)
