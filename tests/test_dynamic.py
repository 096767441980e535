"""The dynamic frame on the GPU against the oracle: per-frame vertex jitter,
full LBVH rebuild, primary rays, two diffuse bounces and a shadow ray per
hit, eagerly and as a captured graph replayed frame after frame."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd import scenes
from rta_amd.abi import TRI_DTYPE, call, ptr
from rta_amd.device import DeviceBuffer
from rta_amd.frame import DynamicFrame, bounce_seed

pytestmark = pytest.mark.gpu
F32MAX = np.float32(np.finfo(np.float32).max)
SEED = 0x5EED


def say(msg):
    print(f"[dynamic] {msg}", flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _records_equal(got, exp):
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(got[f]), bits(exp[f])), f
    assert np.array_equal(got["ref"], exp["ref"])


def _small_frame(bounces=2, shadows=True, concurrent=True):
    v, i = scenes.instances(16, 32, seed=4)
    cam = scenes.dynamic_camera(256, 144, n_inst=16)
    light = scenes.dynamic_light(16)
    fr = DynamicFrame(v, i, cam, light, scenes.DYNAMIC_SIGMA, SEED, bounces=bounces,
                      shadows=shadows, concurrent_shadows=concurrent)
    fr.prime()
    return fr, v, i, cam, light


def test_jitter_bit_exact(gpu):
    say("jitter")
    v, i = scenes.instances(9, 16, seed=2)
    n = len(i) // 3
    vb, ib = DeviceBuffer.from_array(v), DeviceBuffer.from_array(i)
    out = DeviceBuffer(104 * n)
    word = DeviceBuffer(4)
    for frame in (0, 7):
        call("rta_jitter_triangles", ptr(out), ptr(vb), ptr(ib), n, 1e-3, SEED, frame, None, None)
        got = out.download(TRI_DTYPE, n)
        exp = oracle.jitter_triangles(v, i, 1e-3, SEED, frame)
        assert np.array_equal(bits(got), bits(exp)), frame
        # the same frame through the device counter
        word.fill_u32(frame)
        call("rta_jitter_triangles", ptr(out), ptr(vb), ptr(ib), n, 1e-3, SEED, 999, ptr(word),
             None)
        assert np.array_equal(bits(out.download(TRI_DTYPE, n)), bits(exp)), frame


def _oracle_frame(v, i, cam, light, frame, bounces=2):
    tris = oracle.jitter_triangles(v, i, scenes.DYNAMIC_SIGMA, SEED, frame)
    ob = oracle.lbvh_build(tris)
    o, d, t = cam.host_rays()
    hits, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], o, d, t)
    out = {"ob": ob, "hits": [hits], "shadow": []}
    for k in range(bounces + 1):
        so, sd, st, _ = oracle.shadow_rays(hits, ob["tris"], o, d, np.float32(light))
        sh, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], so, sd, st, any_hit=True)
        out["shadow"].append(sh)
        if k < bounces:
            bo, bd, bt, _ = oracle.diffuse_bounce(hits, ob["tris"], o, d,
                                                  bounce_seed(SEED, k + 1))
            h2, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], bo, bd, bt)
            # pad to the full ray count: rays past the bounce count are misses
            full = np.zeros(len(t), h2.dtype)
            full["t"] = F32MAX
            full[:len(h2)] = h2
            hits, o, d = full, bo, bd
            out["hits"].append(full)
    return out


def _compare(res, exp):
    ob = exp["ob"]
    assert np.array_equal(bits(res["tris"]), bits(ob["tris"]))
    assert np.array_equal(res["nodes"]["type_left_elems"], ob["nodes"]["type_left_elems"])
    assert np.array_equal(res["nodes"]["right_tris"], ob["nodes"]["right_tris"])
    assert np.array_equal(res["nodes"]["min"], ob["nodes"]["min"])
    assert np.array_equal(res["nodes"]["max"], ob["nodes"]["max"])
    for k, (g, e) in enumerate(zip(res["hits"], exp["hits"])):
        hit = e["t"] != F32MAX
        assert hit.sum() > 1000 or k > 0
        assert np.array_equal(g["t"] != F32MAX, hit), f"level {k}"
        _records_equal(g[hit], e[hit])
    for k, (g, e) in enumerate(zip(res["shadow_hits"], exp["shadow"])):
        n = res["counts"][f"shadow{k}"]
        assert n == len(e)
        hit = e["t"] != F32MAX
        assert np.array_equal(g["t"][:n] != F32MAX, hit), f"shadow level {k}"
        _records_equal(g[:n][hit], e[hit])


@pytest.mark.parametrize("concurrent", [True, False])
def test_dynamic_frame_eager_bit_exact(gpu, concurrent):
    """Shadow rays on a second stream beside the next bounce, or in
    sequence on the frame's stream: the same records."""
    say(f"eager frame 3, concurrent shadows {concurrent}")
    fr, v, i, cam, light = _small_frame(concurrent=concurrent)
    fr.enqueue(frame=3)
    fr.stream.synchronize()
    fr.check_errors()
    res = fr.results()
    c = res["counts"]
    assert c["bounce1"] > 1000 and c["bounce2"] > 500 and c["shadow0"] > 1000
    _compare(res, _oracle_frame(v, i, cam, light, 3))


def test_dynamic_frame_graph_replays(gpu):
    """A captured frame replayed three times advances its own frame counter
    and reproduces the eager frames 0, 1, 2 bit for bit, with eager work and
    synchronisation between the replays."""
    say("graph replays")
    fr, v, i, cam, light = _small_frame()
    eager = []
    for f in range(3):
        fr.enqueue(frame=f)
        fr.stream.synchronize()
        eager.append(fr.results())
    graph = fr.capture()
    fr.set_frame_counter(0)
    fr.stream.synchronize()
    for f in range(3):
        graph.launch(fr.stream)
        fr.stream.synchronize()
        fr.check_errors()
        got = fr.results()
        assert int(fr.frame_word.download(np.uint32, 1)[0]) == f + 1
        e = eager[f]
        assert np.array_equal(bits(got["tris"]), bits(e["tris"]))
        assert np.array_equal(bits(got["nodes"]), bits(e["nodes"]))
        assert got["counts"] == e["counts"]
        for g, x in zip(got["hits"] + got["shadow_hits"], e["hits"] + e["shadow_hits"]):
            # a miss keeps the stale ref of whatever the buffer held before
            # (the reset writes t, beta, gamma only, cuda-kernels.cu:26-28)
            hit = x["t"] != F32MAX
            assert np.array_equal(g["t"] != F32MAX, hit)
            for f in ("t", "beta", "gamma"):
                assert np.array_equal(bits(g[f]), bits(x[f]))
            assert np.array_equal(g["ref"][hit], x["ref"][hit])
        # unrelated eager work between replays
        rta_amd.synchronize()
    del graph


def test_dynamic_frame_online_finetune(gpu):
    """A hit predictor fine-tuned inside the frame on each frame's first-
    bounce hits: the frame's records stay the oracle's, the loss falls over
    the frames, and the captured frame (fine-tune included) replays with its
    own step and frame counters advancing."""
    from rta_amd.frame import OnlineFineTune
    from rta_amd.predictor import HitPredictor

    say("online fine-tune")
    fr, v, i, cam, light = _small_frame()
    pts = v.reshape(-1, 3)
    pred = HitPredictor(2048, seed=5, scene_min=pts.min(axis=0) - 1.0,
                        scene_max=pts.max(axis=0) + 1.0)
    ft = OnlineFineTune(pred, fr.n, batch=2048, steps=8, lr=2e-3, level=1)
    fr.attach_finetune(ft)
    losses, accs = [], []
    for f in range(24):
        fr.enqueue(frame=f)
        fr.stream.synchronize()
        losses.append(ft.mean_loss())
        accs.append(ft.accuracy())
        if f == 3:
            fr.check_errors()
            _compare(fr.results(), _oracle_frame(v, i, cam, light, 3))
    say(f"loss {losses[0]:.3f} -> {losses[-1]:.3f}, top-4 accuracy before the frame's steps "
        f"{accs[0]:.3f} -> {np.mean(accs[-4:]):.3f}")
    assert int(ft.step.download(np.uint32, 1)[0]) == 24 * 8
    assert np.mean(losses[-4:]) < 0.8 * losses[0]
    assert np.mean(accs[-4:]) > 5 * max(accs[0], 4 / 1024)
    graph = fr.capture()
    fr.set_frame_counter(24)
    fr.stream.synchronize()
    before = pred.params()
    for r in range(3):
        graph.launch(fr.stream)
        fr.stream.synchronize()
        fr.check_errors()
        assert int(fr.frame_word.download(np.uint32, 1)[0]) == 25 + r
        assert int(ft.step.download(np.uint32, 1)[0]) == 24 * 8 + 8 * (r + 1)
    assert not np.array_equal(pred.params(), before)
    assert np.isfinite(ft.mean_loss())
    del graph
