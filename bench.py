#!/usr/bin/env python3
"""Benchmark of the GPU BVH path on the 2.3M-triangle workload.

Workload (one step = one frame): 1024x1024 heightfield (6-octave value noise,
amplitude 0.2) + 200,000 random small triangles = 2,297,152 triangles;
full LBVH build (7 stages) + child-pair packing, 1920x1080 pinhole primary
rays traced, one cosine-weighted diffuse bounce per hit generated on the GPU
and traced (the incoherent ray set).  All inputs are resident in HBM before
the timed region; the pristine triangle array is restored by a device copy at
the start of each step (the build reorders it in place, as the reference).

`value` is the plain closest-hit Mrays/s on the bounce rays (kernel time from
HIP events on the stream the kernels run on).  Primary-ray Mrays/s and LBVH
build Mtri/s ride along in the same line, with the hit-predicted leg on the
same scene (`hit_predicted`), the 256-mesh batch (`batch`) and the 8.4 M
triangle dynamic frame (`dynamic`), each measured in this run
(`--workload batch|dynamic` prints either as a line of its own).  The CPU
baseline is the oracle's restatement of the reference CPU bbvh plugin
(spatial-median tree, child-is tracer, OpenMP over rows) on the same bounce
rays, timed on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "closest-hit Mrays/s (plain vs hit-predicted) at 2M tris; LBVH build Mtri/s"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--grid", type=int, default=1024)
    p.add_argument("--random", type=int, default=200_000)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--trav", choices=("persistent", "packed"), default="persistent")
    p.add_argument("--block", type=int, default=64)
    p.add_argument("--refill", type=int, default=16)
    p.add_argument("--steps-per-refill", dest="steps_per_refill", type=int, default=32)
    p.add_argument("--leaf-min", dest="leaf_min", type=int, default=10)
    p.add_argument("--primary-stack", dest="primary_stack", type=int, default=16,
                   choices=(16, 24, 32),
                   help="LDS stack entries of the primary trace (rays that need more are "
                        "traced again with 32); the bounce trace keeps 32")
    p.add_argument("--serial-shadows", dest="serial_shadows", action="store_true",
                   help="dynamic frame: trace each level's shadow rays on the frame's stream "
                        "instead of a second stream beside the next bounce")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-legs", dest="no_legs", action="store_true",
                   help="skip the many-mesh and dynamic-frame legs of the default run")
    p.add_argument("--no-predictor", action="store_true",
                   help="skip the hit-predicted leg (training + seeded traversal)")
    p.add_argument("--predictor-epochs", dest="predictor_epochs", type=int, default=8)
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a 1 GiB device copy kernel (to scale HBM traffic counters)")
    p.add_argument("--cpu-rays", type=int, default=0,
                   help="bounce rays in the CPU sample (0 = all of them)")
    p.add_argument("--workload", choices=("terrain", "dynamic", "batch"), default="terrain",
                   help="terrain: the 2.3M-triangle heightfield scene (default); dynamic: "
                        "64 instances x 131,072 triangles rebuilt every frame, 3840x2160, "
                        "2 bounces + shadow rays, frames replayed from a captured graph; "
                        "batch: 256 meshes built in shared launches and traced in one "
                        "persistent launch")
    p.add_argument("--instances", type=int, default=64)
    p.add_argument("--instance-grid", dest="instance_grid", type=int, default=256)
    return p.parse_args()


def batch_leg(args) -> dict:
    """256 meshes (1k-200k triangles, 4k-256k rays each): all builds in one
    launch per stage, all rays in one autotuned persistent launch; beside
    it the same work as one build and one trace launch per mesh."""
    import ctypes as C

    import rta_amd
    from rta_amd import scenes
    from rta_amd.abi import NODE_DTYPE, TraceConfig, call, ptr
    from rta_amd.batch import MeshBatch
    from rta_amd.device import DeviceBuffer, Event

    t0 = time.perf_counter()
    meshes, rays = scenes.batch_meshes(256)
    log(f"batch: {len(meshes)} meshes, {sum(len(m) for m in meshes)} triangles, "
        f"{sum(len(r[2]) for r in rays)} rays ({time.perf_counter() - t0:.1f} s)")
    b = MeshBatch(meshes, rays)
    s = b.stream
    a, z = Event(), Event()

    def timed(fn, reps):
        best = 1e30
        for _ in range(reps):
            a.record(s)
            fn()
            z.record(s)
            z.synchronize()
            best = min(best, a.elapsed_ms(z))
        return best

    b.build()
    s.synchronize()
    build_ms = timed(b.build, max(args.steps, 3))
    cfg, tune_ms = b.autotune(rounds=2, reps=2)
    log(f"autotuned {cfg.as_tuple()} -> {tune_ms:.3f} ms")
    trace_ms = timed(lambda: b.trace(config=cfg), max(args.steps, 3))
    ovf = b.overflow_count()
    default_ms = timed(lambda: b.trace(config=TraceConfig(64, 16, 32, 0, 10)), 3)
    b.check_errors()
    # the whole frame (restore + build + pack + reset + trace) as one graph
    g = b.capture(config=cfg)
    graph_ms = timed(lambda: g.launch(s), max(args.steps, 3))
    b.check_errors()
    del g

    # the same meshes one build + one trace launch at a time
    per = []
    for m, mesh in enumerate(meshes):
        n = len(mesh)
        per.append({
            "src": DeviceBuffer(104 * n), "nodes": DeviceBuffer(32 * (2 * n - 1)),
            "idx": DeviceBuffer(4 * n), "pairs": DeviceBuffer(64 * (2 * n - 1)),
            "tri4": DeviceBuffer(48 * n), "flags": DeviceBuffer(4), "n": n,
            "r0": int(b.ray_off[m]), "nr": int(b.n_rays[m]),
        })
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", max(len(x) for x in meshes), C.byref(sz))
    ws = DeviceBuffer(sz.value)
    rta_amd.synchronize()
    sp = ptr(s)

    def build_each():
        for m, q in enumerate(per):
            call("rta_hip_memcpy_dtod", ptr(q["src"]), C.c_void_p(b.pristine.ptr + 104 * int(b.tri_off[m])),
                 C.c_size_t(104 * q["n"]), sp)
            call("rta_lbvh_build_packed", ptr(q["src"]), q["n"], ptr(q["nodes"]), ptr(q["idx"]),
                 ptr(q["pairs"]), ptr(q["tri4"]), ptr(q["flags"]), ptr(ws), C.c_size_t(sz.value),
                 None, None, sp)

    build_each()
    s.synchronize()
    roots = []
    for q in per:
        r = q["nodes"].download(NODE_DTYPE, 1)[0]
        roots.append((int(r["type_left_elems"]), int(r["right_tris"])))

    # per-mesh launches: the autotuned shape with the full stack (a short
    # stack needs the batch's overflow list)
    cfg_each = TraceConfig(*cfg.as_tuple()[:5], 32)

    def trace_each():
        b.reset_hits()
        for m, q in enumerate(per):
            if q["nr"] == 0:
                continue
            r0 = q["r0"]
            call("rta_trace_cis_persistent", ptr(q["tri4"]), ptr(q["pairs"]), ptr(q["flags"]),
                 roots[m][0], roots[m][1], C.c_void_p(b.ro.ptr + 12 * r0),
                 C.c_void_p(b.rd.ptr + 12 * r0), C.c_void_p(b.rt.ptr + 4 * r0), q["nr"], None,
                 C.c_void_p(b.hits.ptr + 16 * r0), ptr(b.err), ptr(b.queue), C.byref(cfg_each), sp)

    each_build_ms = timed(build_each, 3)
    each_trace_ms = timed(trace_each, 3)
    R, N = b.total_rays, b.total_tris
    log(f"batch: build {build_ms:.3f} ms, trace {trace_ms:.3f} ms ({R / (trace_ms * 1e3):.1f} "
        f"Mrays/s); one launch per mesh: build {each_build_ms:.3f} ms, trace "
        f"{each_trace_ms:.3f} ms")
    return {
        "workload": ("256 independent meshes (1k-200k triangles, jittered icospheres and "
                     "heightfield patches) with 4k-256k rays each (half camera, half "
                     "random): all LBVH builds in one launch per stage, all rays in one "
                     "persistent launch with an autotuned shape"),
        "meshes": len(meshes), "n_tris": N, "rays": R,
        "batch_mrays_s": round(R / (trace_ms * 1e3), 3),
        "steps": max(args.steps, 3),
        "batch_build_ms": round(build_ms, 4),
        "batch_build_mtri_s": round(N / (build_ms * 1e3), 3),
        "batch_trace_ms": round(trace_ms, 4),
        "autotuned_config": list(cfg.as_tuple()),
        "autotuned_config_fields": ["block", "refill_min", "steps", "blocks_per_cu",
                                    "leaf_min", "stack"],
        "short_stack_rays_retraced": ovf,
        "graph_frame_ms": round(graph_ms, 4),
        "graph_frame": "restore + build with the traversal layout + reset + trace captured as one graph",
        "default_config_trace_ms": round(default_ms, 4),
        "per_mesh_launches_build_ms": round(each_build_ms, 4),
        "per_mesh_launches_trace_ms": round(each_trace_ms, 4),
        "batching_speedup_build": round(each_build_ms / build_ms, 3),
        "batching_speedup_trace": round(each_trace_ms / trace_ms, 3),
    }


def batch_main(args) -> None:
    import rta_amd

    leg = batch_leg(args)
    out = {
        "metric": METRIC,
        "value": leg["batch_mrays_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": leg["steps"],
        "warmup": 1,
        "ms_per_step": round(leg["batch_build_ms"] + leg["batch_trace_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": leg["workload"] + "; value = batched Mrays/s",
                   "meshes": leg["meshes"], "n_tris": leg["n_tris"], "rays": leg["rays"]},
        **{k: v for k, v in leg.items() if k not in ("workload", "meshes", "n_tris", "rays",
                                                     "batch_mrays_s", "steps")},
        "device": rta_amd.device_name(),
    }
    print(json.dumps(out), flush=True)


def dynamic_leg(args) -> dict:
    """The dynamic frame: per-frame jitter, full rebuild, primary rays, two
    diffuse bounces and a shadow ray per hit, replayed from a graph."""
    import rta_amd
    from rta_amd import scenes
    from rta_amd.abi import TraceConfig
    from rta_amd.device import Event
    from rta_amd.frame import DynamicFrame

    t0 = time.perf_counter()
    v, i = scenes.instances(args.instances, args.instance_grid, seed=scenes.DYNAMIC_SEED)
    n = len(i) // 3
    w, h = (args.width, args.height) if (args.width, args.height) != (1920, 1080) else (3840, 2160)
    cam = scenes.dynamic_camera(w, h, n_inst=args.instances)
    log(f"dynamic scene {n} triangles, {len(v) // 3} vertices, {w}x{h} "
        f"({time.perf_counter() - t0:.1f} s)")
    cfg = TraceConfig(args.block, args.refill, args.steps_per_refill, 0, args.leaf_min)
    fr = DynamicFrame(v, i, cam, scenes.dynamic_light(args.instances), scenes.DYNAMIC_SIGMA,
                      scenes.DYNAMIC_SEED, bounces=2, shadows=True, config=cfg,
                      concurrent_shadows=not args.serial_shadows)
    fr.prime()
    # eager frames with stage events (stages in sequence): the per-stage breakdown
    stage_names = ["jitter", "build+pack", "primary", "shadow0", "bounce1", "shadow1", "bounce2",
                   "shadow2"]
    stage_ms = {k: 0.0 for k in stage_names}
    n_eager = max(1, args.warmup)
    for f in range(n_eager):
        fr.enqueue(frame=f, timed=True)
        fr.stream.synchronize()
        ev = fr.events
        for k, name in enumerate(stage_names):
            stage_ms[name] += ev[k].elapsed_ms(ev[k + 1]) / n_eager
    fr.check_errors()
    counts = fr.ray_counts()
    graph = fr.capture()
    fr.set_frame_counter(n_eager)
    fr.stream.synchronize()
    graph.launch(fr.stream)                     # first replay (untimed)
    fr.stream.synchronize()
    a, b = Event(), Event()
    call_sync = rta_amd.synchronize
    call_sync()
    t_start = time.perf_counter()
    a.record(fr.stream)
    for _ in range(args.steps):
        graph.launch(fr.stream)
    b.record(fr.stream)
    b.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    frame_ms = a.elapsed_ms(b) / args.steps
    fr.check_errors()
    counts_last = fr.ray_counts()
    rays = sum(counts_last.values())
    del graph
    log(f"dynamic: {frame_ms:.3f} ms per captured frame, {rays / (frame_ms * 1e3):.1f} Mrays/s "
        f"over the frame's {rays} rays")
    finetune = None if args.no_predictor else dynamic_finetune(args, fr, v, n_eager)
    return {
        "workload": (f"dynamic scene: {args.instances} instances x "
                     f"{2 * args.instance_grid ** 2} triangles = {n}, per-vertex jitter "
                     f"sigma 1e-3 and full LBVH rebuild every frame, {w}x{h} primary rays, "
                     "2 diffuse bounces and 1 shadow ray per hit, frames replayed from one "
                     "captured graph"),
        "n_tris": n, "rays_per_frame": counts_last,
        "primary_stack": fr.primary_config.stack,
        "concurrent_shadows": fr.concurrent,
        "frame_mrays_s": round(rays / (frame_ms * 1e3), 3),
        "steps": args.steps, "warmup": n_eager,
        "frame_wall_ms": round(wall_ms, 4),
        "frame_ms": round(frame_ms, 4),
        "frames_per_s": round(1e3 / frame_ms, 2),
        "build_and_pack_mtri_s": round(n / (stage_ms["build+pack"] * 1e3), 3),
        "eager_stage_ms": {k: round(x, 4) for k, x in stage_ms.items()},
        "eager_counts_first_frames": counts,
        "online_finetune": finetune,
    }


def dynamic_main(args) -> None:
    import rta_amd

    leg = dynamic_leg(args)
    out = {
        "metric": METRIC,
        "value": leg["frame_mrays_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": leg["steps"],
        "warmup": leg["warmup"],
        "ms_per_step": leg["frame_wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": leg["workload"] + "; value = all rays traced per frame / frame "
                                                 "time",
                   "n_tris": leg["n_tris"], "rays_per_frame": leg["rays_per_frame"]},
        **{k: v for k, v in leg.items() if k not in ("workload", "n_tris", "rays_per_frame",
                                                     "frame_mrays_s", "steps", "warmup",
                                                     "frame_wall_ms")},
        "device": rta_amd.device_name(),
    }
    print(json.dumps(out), flush=True)


def dynamic_finetune(args, fr, v, first_frame):
    """The dynamic frame with a hit predictor fine-tuned online on each
    frame's first-bounce hits (4 Adam steps on 8,192 sampled rays per frame,
    from an untrained network): loss and top-4 accuracy of the predictions
    made before each frame's steps over eager frames, the stage's time, and
    the frame time of the captured frame with the stage inside."""
    from rta_amd.device import Event
    from rta_amd.frame import OnlineFineTune
    from rta_amd.predictor import HitPredictor

    pts = v.reshape(-1, 3)
    pred = HitPredictor(8192, seed=1, scene_min=pts.min(axis=0) - 1.0,
                        scene_max=pts.max(axis=0) + 1.0)
    ft = OnlineFineTune(pred, fr.n, batch=8192, steps=4, lr=2e-3, level=1)
    fr.attach_finetune(ft)
    names = ["jitter", "build+pack", "primary", "shadow0", "bounce1", "finetune", "shadow1",
             "bounce2", "shadow2"]
    frames = max(40, args.warmup)
    losses, accs, ft_ms = [], [], []
    for f in range(frames):
        fr.enqueue(frame=first_frame + f, timed=True)
        fr.stream.synchronize()
        ev = fr.events
        ft_ms.append(ev[names.index("finetune")].elapsed_ms(ev[names.index("finetune") + 1]))
        losses.append(ft.mean_loss())
        accs.append(ft.accuracy())
    fr.check_errors()
    graph = fr.capture()
    fr.set_frame_counter(first_frame + frames)
    fr.stream.synchronize()
    graph.launch(fr.stream)
    fr.stream.synchronize()
    a, b = Event(), Event()
    a.record(fr.stream)
    for _ in range(args.steps):
        graph.launch(fr.stream)
    b.record(fr.stream)
    b.synchronize()
    fr.check_errors()
    frame_ms = a.elapsed_ms(b) / args.steps
    fr.attach_finetune(None)
    del graph
    log(f"online fine-tune: {np.median(ft_ms):.3f} ms per frame, loss {losses[0]:.3f} -> "
        f"{losses[-1]:.3f}, top-4 accuracy {accs[0]:.3f} -> {accs[-1]:.3f}, frame with it "
        f"{frame_ms:.3f} ms")
    marks = sorted({0, 1, 2, 5, 10, 20, frames - 1})
    return {
        "batch": 8192, "steps_per_frame": 4, "lr": 2e-3, "level": "bounce1", "eager_frames": frames,
        "stage_ms_median": round(float(np.median(ft_ms)), 4),
        "frame_ms_with_finetune": round(frame_ms, 4),
        "loss_by_frame": {str(k): round(losses[k], 4) for k in marks},
        "top4_accuracy_by_frame": {str(k): round(accs[k], 4) for k in marks},
        "note": ("accuracy of the predictions made before each frame's own steps, on that "
                 "frame's first sample; the treelet cut is recomputed from every frame's "
                 "rebuilt tree"),
    }


def cpu_baseline(tris_h, bo, bd, bt, n_bounce, w, max_rays):
    """The oracle restatement of the reference CPU plugin on this host."""
    import oracle

    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS"):
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    t0 = time.perf_counter()
    mb = oracle.median_build(tris_h)
    build_s = time.perf_counter() - t0
    # the CPU reference point for the LBVH build: the oracle's single-threaded
    # restatement of the seven lbvh-kernels.cu stages on the same triangles
    t0 = time.perf_counter()
    oracle.lbvh_build(tris_h)
    lbvh_s = time.perf_counter() - t0
    rows = n_bounce // w
    if max_rays:
        rows = min(rows, max(1, max_rays // w))
    n = rows * w
    hits, ms, _ = oracle.trace_cis(mb["tris"], mb["nodes"], bo[: 3 * n], bd[: 3 * n], bt[:n], w,
                                   rows, threads=threads)
    cpu = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": n / ms / 1e3, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": (f"{n} of the {n_bounce} bounce rays ({rows} rows of {w}) traced by the "
                   f"oracle's child-is tracer on its spatial-median tree (built single-threaded "
                   f"in {build_s:.1f} s, not timed), OpenMP schedule(dynamic,32), {cpu}"),
        "median_build_s": build_s,
        "lbvh_build_s": lbvh_s,
        "lbvh_build_mtri_s": len(tris_h) / lbvh_s / 1e6,
        "lbvh_build_threads": 1,
    }


PRED_FLOPS_PER_RAY = 2 * (102 * 64 + 64 * 64 + 64 * 64 + 64 * 1025)   # 160,640
# a training step per ray: the forward pass, the data gradients of layers
# 4, 3, 2 and the weight gradients of all four (468,864)
PRED_TRAIN_FLOPS_PER_RAY = 2 * PRED_FLOPS_PER_RAY + 2 * (64 * 1025 + 64 * 64 + 64 * 64)


def predictor_leg(args, tris_h, tris, nodes, packed, ranges, prim, hits_p, R, tcfg, s):
    """Plain against hit-predicted closest-hit traversal of bounce rays.

    The network is trained on the diffuse bounce rays of four seeds and
    evaluated on two held-out seeds (the first is the headline bounce set).
    The classes are a fixed cut of the tree into 1024 treelets.
    Hit-predicted = inference (top-4 treelets, fused kernel), the traversal
    of those treelets (a max_t seed), then the unchanged persistent
    traversal with max_t' = min(max_t, seed); all three inside the timing.
    Every held-out ray's record is compared with the plain traversal's."""
    import rta_amd
    from rta_amd import predictor as P
    from rta_amd import scenes
    from rta_amd.abi import HIT_DTYPE, call, ptr
    from rta_amd.device import DeviceBuffer, Event

    n = len(tris_h)
    sp = ptr(s)
    sz = C.c_size_t(0)
    call("rta_bounce_work_bytes", R, C.byref(sz))
    bwork, bwork_bytes = DeviceBuffer(sz.value), C.c_size_t(sz.value)
    src = DeviceBuffer(4 * R)
    err = DeviceBuffer(4)
    err.fill_u32(0)
    queue = DeviceBuffer(4)
    fmax = np.float32(rta_amd.FLT_MAX)

    def persistent(rays, cnt, maxt, hits):
        call("rta_trace_cis_persistent", ptr(packed.tri4), ptr(packed.pairs), ptr(packed.flags),
             packed.root_link, packed.root_index, ptr(rays.gpu_origin), ptr(rays.gpu_direction),
             ptr(maxt), R, ptr(cnt), ptr(hits), ptr(err), ptr(queue), C.byref(tcfg), sp)

    def bounce_set(seed, rays, cnt, hits):
        call("rta_diffuse_bounce", ptr(rays.gpu_origin), ptr(rays.gpu_direction),
             ptr(rays.gpu_maxt), ptr(src), ptr(cnt), ptr(hits_p), ptr(tris), ptr(prim.gpu_origin),
             ptr(prim.gpu_direction), R, C.c_uint64(seed), ptr(bwork), bwork_bytes, sp)
        call("rta_reset_intersections", ptr(hits), R, 1, sp)
        persistent(rays, cnt, rays.gpu_maxt, hits)
        s.synchronize()
        return int(cnt.download(np.uint32, 1)[0])

    # the fixed cut into 1024 treelets (the network's classes)
    t0 = time.perf_counter()
    warm = P.TreeletCut(packed, ranges, n, stream=s)    # warm-up (first launch loads code)
    ev_cut = [Event(), Event()]
    ev_cut[0].record(s)
    cut = P.TreeletCut(packed, ranges, n, stream=s)
    ev_cut[1].record(s)
    ev_cut[1].synchronize()
    del warm
    cut_ms = ev_cut[0].elapsed_ms(ev_cut[1])
    sizes = np.diff(np.append(cut.host()["first"].astype(np.int64), n))

    # training data: bounce rays of four seeds and their plain records
    tmp, tcnt, thits = rta_amd.RayBuffers(R), DeviceBuffer(4), DeviceBuffer(16 * R)
    o_tr, d_tr, lab_tr, tgt_tr = [], [], [], []
    for k in scenes.PREDICTOR_TRAIN_SEEDS:
        m = bounce_set(scenes.BOUNCE_SEED + k, tmp, tcnt, thits)
        o, d, _ = tmp.download(m)
        lab, tgt = cut.labels_host(thits, m, stream=s)
        o_tr.append(o), d_tr.append(d), lab_tr.append(lab), tgt_tr.append(tgt)
    o_tr, d_tr = np.concatenate(o_tr), np.concatenate(d_tr)
    lab_tr, tgt_tr = np.concatenate(lab_tr), np.concatenate(tgt_tr)
    n_tr = len(lab_tr)
    del tmp, thits

    evals = []
    for k in scenes.PREDICTOR_EVAL_SEEDS:
        e = {"seed": scenes.BOUNCE_SEED + k, "rays": rta_amd.RayBuffers(R), "cnt": DeviceBuffer(4),
             "plain": DeviceBuffer(16 * R), "pred": DeviceBuffer(16 * R)}
        e["m"] = bounce_set(e["seed"], e["rays"], e["cnt"], e["plain"])
        evals.append(e)
    m_max = max(e["m"] for e in evals)
    log(f"predictor: {n_tr} training rays, held-out {[e['m'] for e in evals]} "
        f"({time.perf_counter() - t0:.1f} s)")

    allv = np.concatenate([tris_h["a"], tris_h["b"], tris_h["c"]])
    pred = P.HitPredictor(8192, seed=1, scene_min=allv.min(axis=0), scene_max=allv.max(axis=0))
    o_b, d_b = DeviceBuffer.from_array(o_tr), DeviceBuffer.from_array(d_tr)
    l_b, g_b = DeviceBuffer.from_array(lab_tr), DeviceBuffer.from_array(tgt_tr)
    losses, train_s = pred.train(o_b, d_b, l_b, g_b, n_tr, epochs=args.predictor_epochs,
                                 batch=8192, lr=2e-3, stream=s)
    del o_b, d_b, l_b, g_b
    train_steps = args.predictor_epochs * ((n_tr + 8191) // 8192)
    log(f"predictor trained: {args.predictor_epochs} epochs ({train_steps} steps) in "
        f"{train_s:.2f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}")

    top, t_seed, maxt2 = DeviceBuffer(16 * m_max), DeviceBuffer(4 * m_max), DeviceBuffer(4 * m_max)
    ev = [Event() for _ in range(5)]
    reps = max(args.steps, 3)
    out = []
    for e in evals:
        m, rays = e["m"], e["rays"]
        # plain: the same persistent traversal as the headline figure
        plain_ms = 0.0
        for _ in range(reps):
            call("rta_reset_intersections", ptr(e["plain"]), R, 1, sp)
            ev[0].record(s)
            persistent(rays, e["cnt"], rays.gpu_maxt, e["plain"])
            ev[1].record(s)
            ev[1].synchronize()
            plain_ms += ev[0].elapsed_ms(ev[1]) / reps
        hp = e["plain"].download(HIT_DTYPE, m)
        hit = hp["t"] != fmax
        lab, _ = cut.labels_host(e["plain"], m, stream=s)
        stats = DeviceBuffer(8 * m)
        scratch = DeviceBuffer(16 * m)

        def visit_counts(mt):
            """per-ray inner-node visits and triangle tests with max_t mt"""
            call("rta_reset_intersections", ptr(scratch), m, 1, sp)
            call("rta_trace_cis", ptr(tris), n, ptr(nodes), ptr(rays.gpu_origin),
                 ptr(rays.gpu_direction), ptr(mt), m, 1, ptr(scratch), ptr(stats), ptr(err), sp)
            s.synchronize()
            return stats.download(np.uint32, 2 * m).reshape(-1, 2).astype(np.float64)

        def visits(mt):
            """mean inner-node visits and triangle tests per ray with max_t mt"""
            vt = visit_counts(mt)
            return round(float(vt[:, 0].mean()), 3), round(float(vt[:, 1].mean()), 3)

        rec = {"bounce_seed": e["seed"], "rays": m,
               "plain_mrays_s": round(m / (plain_ms * 1e3), 3), "plain_ms": round(plain_ms, 4),
               "hit_share": round(float(hit.mean()), 4), "by_candidates": []}
        vt_plain = visit_counts(rays.gpu_maxt)
        rec["V_plain"] = round(float(vt_plain[:, 0].mean()), 3)
        rec["T_plain"] = round(float(vt_plain[:, 1].mean()), 3)
        # The ceiling of any bound seeding: the perfect seed, each hit ray's
        # own closest t (widened by the kernels' seed margin), as max_t.  It
        # changes no record; the visits it leaves bound what a predictor can
        # save, and V_plain / V_perfect_seed bounds the speed-up of a
        # traversal whose time follows its inner-node visits.
        t_best = DeviceBuffer.from_array(np.where(hit, hp["t"], fmax).astype(np.float32))
        maxt_c = DeviceBuffer(4 * m)
        call("rta_seed_max_t", ptr(maxt_c), ptr(rays.gpu_maxt), ptr(t_best), m, sp)
        vt_c = visit_counts(maxt_c)
        rec["ceiling"] = {
            "V_perfect_seed": round(float(vt_c[:, 0].mean()), 3),
            "T_perfect_seed": round(float(vt_c[:, 1].mean()), 3),
            "V_plain_hit_rays": round(float(vt_plain[hit, 0].mean()), 3),
            "V_perfect_seed_hit_rays": round(float(vt_c[hit, 0].mean()), 3),
            "visit_saving": round(1.0 - float(vt_c[:, 0].mean() / vt_plain[:, 0].mean()), 4),
            "speedup_bound": round(float(vt_plain[:, 0].mean() / vt_c[:, 0].mean()), 4),
        }
        del t_best, maxt_c
        log(f"predictor seed {e['seed']}: perfect-seed ceiling V {rec['V_plain']:.2f} -> "
            f"{rec['ceiling']['V_perfect_seed']:.2f} (speed-up bound "
            f"{rec['ceiling']['speedup_bound']:.3f})")
        # the seeding traverses the first 1, 2 or all 4 proposed treelets
        for cand in (1, 2, 4):
            part = {"inference": 0.0, "seeding": 0.0, "traversal": 0.0}
            total_ms = 0.0
            for _ in range(reps):
                call("rta_reset_intersections", ptr(e["pred"]), R, 1, sp)
                ev[0].record(s)
                pred.infer(rays.gpu_origin, rays.gpu_direction, m, top=top, stream=s,
                           top_k=1 if cand == 1 else 4)
                ev[1].record(s)
                call("rta_trace_seed_treelets", ptr(packed.tri4), ptr(packed.pairs),
                     ptr(packed.flags), ptr(cut.table), ptr(cut.count), ptr(rays.gpu_origin),
                     ptr(rays.gpu_direction), ptr(rays.gpu_maxt), m, ptr(top), cand,
                     ptr(t_seed), ptr(err), sp)
                call("rta_seed_max_t", ptr(maxt2), ptr(rays.gpu_maxt), ptr(t_seed), m, sp)
                ev[2].record(s)
                persistent(rays, e["cnt"], maxt2, e["pred"])
                ev[3].record(s)
                ev[3].synchronize()
                part["inference"] += ev[0].elapsed_ms(ev[1]) / reps
                part["seeding"] += ev[1].elapsed_ms(ev[2]) / reps
                part["traversal"] += ev[2].elapsed_ms(ev[3]) / reps
                total_ms += ev[0].elapsed_ms(ev[3]) / reps
            if int(err.download(np.uint32, 1)[0]):
                raise SystemExit("hit-predicted leg: traversal reported stack/guard errors")
            hq = e["pred"].download(HIT_DTYPE, m)
            same = (hq["t"] != fmax) == hit
            for f in ("t", "beta", "gamma"):
                same &= (hq[f].view(np.uint32) == hp[f].view(np.uint32)) | ~hit
            same &= (hq["ref"] == hp["ref"]) | ~hit
            seeds = t_seed.download(np.float32, m)
            tops = top.download(np.int32, 4 * m).reshape(m, 4)
            v_s, t_s = visits(maxt2)
            c = {"candidates": cand,
                 "hit_predicted_mrays_s": round(m / (total_ms * 1e3), 3),
                 "hit_predicted_ms": round(total_ms, 4),
                 "breakdown_ms": {k: round(v, 4) for k, v in part.items()},
                 "exact_share": float(same.mean()),
                 "accuracy": round(P.top4_accuracy(tops[:, :cand], lab), 4),
                 "seeded_share": round(float((seeds != fmax).mean()), 4),
                 "seed_is_closest_share": round(float((seeds[hit] == hp["t"][hit]).mean()), 4),
                 "V_seeded": v_s, "T_seeded": t_s}
            rec["by_candidates"].append(c)
            log(f"predictor seed {e['seed']}, {cand} candidate(s): plain "
                f"{rec['plain_mrays_s']:.1f} Mrays/s, hit-predicted "
                f"{c['hit_predicted_mrays_s']:.1f} Mrays/s ({c['breakdown_ms']}), accuracy "
                f"{c['accuracy']:.3f}, exact {c['exact_share']:.6f}, V {rec['V_plain']:.1f} -> "
                f"{v_s:.1f}")
            if c["exact_share"] != 1.0:
                raise SystemExit("hit-predicted leg: seeded records differ from the plain ones")
            if cand == 4:
                rec["top4_accuracy"] = c["accuracy"]
                rec["inference_ms"] = round(part["inference"], 4)
                rec["inference_tflop_s"] = round(
                    PRED_FLOPS_PER_RAY * m / (part["inference"] * 1e-3) / 1e12, 3)
        # top-1 inference + the seeding fused into the persistent traversal
        part = {"inference": 0.0, "seeded_traversal": 0.0}
        total_ms = 0.0
        for _ in range(reps):
            call("rta_reset_intersections", ptr(e["pred"]), R, 1, sp)
            ev[0].record(s)
            pred.infer(rays.gpu_origin, rays.gpu_direction, m, top=top, stream=s, top_k=1)
            ev[1].record(s)
            call("rta_trace_cis_persistent_seeded", ptr(packed.tri4), ptr(packed.pairs),
                 ptr(packed.flags), packed.root_link, packed.root_index, ptr(rays.gpu_origin),
                 ptr(rays.gpu_direction), ptr(rays.gpu_maxt), R, ptr(e["cnt"]), ptr(cut.table),
                 ptr(cut.count), ptr(top), ptr(e["pred"]), ptr(err), ptr(queue), C.byref(tcfg),
                 sp)
            ev[2].record(s)
            ev[2].synchronize()
            part["inference"] += ev[0].elapsed_ms(ev[1]) / reps
            part["seeded_traversal"] += ev[1].elapsed_ms(ev[2]) / reps
            total_ms += ev[0].elapsed_ms(ev[2]) / reps
        if int(err.download(np.uint32, 1)[0]):
            raise SystemExit("hit-predicted leg: traversal reported stack/guard errors")
        hq = e["pred"].download(HIT_DTYPE, m)
        same = (hq["t"] != fmax) == hit
        for f in ("t", "beta", "gamma"):
            same &= (hq[f].view(np.uint32) == hp[f].view(np.uint32)) | ~hit
        same &= (hq["ref"] == hp["ref"]) | ~hit
        c = {"candidates": 1, "fused": True,
             "hit_predicted_mrays_s": round(m / (total_ms * 1e3), 3),
             "hit_predicted_ms": round(total_ms, 4),
             "breakdown_ms": {k: round(v, 4) for k, v in part.items()},
             "exact_share": float(same.mean())}
        rec["by_candidates"].append(c)
        log(f"predictor seed {e['seed']}, fused (top-1 seeding inside the persistent "
            f"traversal): hit-predicted {c['hit_predicted_mrays_s']:.1f} Mrays/s "
            f"({c['breakdown_ms']}), exact {c['exact_share']:.6f}")
        if c["exact_share"] != 1.0:
            raise SystemExit("hit-predicted leg: fused seeded records differ from the plain ones")
        best = max(rec["by_candidates"], key=lambda c: c["hit_predicted_mrays_s"])
        rec["hit_predicted_mrays_s"] = best["hit_predicted_mrays_s"]
        rec["hit_predicted_best_candidates"] = best["candidates"]
        rec["hit_predicted_best_fused"] = bool(best.get("fused", False))
        rec["hit_predicted_vs_plain"] = round(best["hit_predicted_mrays_s"] /
                                              rec["plain_mrays_s"], 4)
        rec["exact_share"] = min(c["exact_share"] for c in rec["by_candidates"])
        out.append(rec)
    pred.close()
    return {
        "ceiling": dict(out[0]["ceiling"], bounce_seed=out[0]["bounce_seed"]),
        "network": "102-64-64-64-1025 MLP, bf16 weights/activations, fp32 MFMA accumulation",
        "flops_per_ray": PRED_FLOPS_PER_RAY,
        "mfma_peak_tflop_s": 2500.0,
        "treelets": {"count": int(len(sizes)), "cut_ms": round(cut_ms, 4),
                     "triangles_min_median_max": [int(sizes.min()), int(np.median(sizes)),
                                                  int(sizes.max())]},
        "training": {"rays": n_tr, "epochs": args.predictor_epochs, "batch": 8192, "lr": 2e-3,
                     "seconds": round(train_s, 3),
                     "path": "fused step: rta_pred_train_batches, 3 launches per batch",
                     "steps": train_steps,
                     "us_per_step": round(train_s / train_steps * 1e6, 1),
                     "flops_per_ray": PRED_TRAIN_FLOPS_PER_RAY,
                     "tflop_s": round(PRED_TRAIN_FLOPS_PER_RAY * n_tr * args.predictor_epochs /
                                      train_s / 1e12, 2),
                     "loss_first_last": [round(losses[0], 4), round(losses[-1], 4)],
                     "bounce_seeds": [scenes.BOUNCE_SEED + k
                                      for k in scenes.PREDICTOR_TRAIN_SEEDS]},
        "held_out": out,
    }


def traversal_roofline(args, traffic, traffic_src, limiter, bounce_ms, b_ray, logical_gbs, V,
                       T):
    """The bounce launch against the HBM roofline.  achieved/frac use the
    HBM bytes the launch really moves (rocprofv3 FETCH_SIZE and WRITE_SIZE
    per dispatch, scaled by a copy kernel of known size, from the committed
    profile of this workload) over this run's mean launch time, so frac <= 1.
    The logical bytes of the per-ray model 28 + 16 + 64 V + 36 T count every
    node and triangle fetch, most of which L2 and the last-level cache serve;
    they are reported beside it and are not an HBM rate."""
    gbs = traffic / (bounce_ms * 1e-3) / 1e9 if traffic else None
    out = {
        "bound": "hbm",
        "achieved": round(gbs, 1) if gbs else None,
        "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None,
        "traffic": traffic, "traffic_source": traffic_src,
        "kernel": f"trace_cis_{args.trav}_kernel (bounce rays)",
        "basis": ("measured HBM bytes per bounce launch (FETCH_SIZE x fetch factor + "
                  "WRITE_SIZE) / this run's mean bounce launch time"),
        "algorithmic": {
            "bytes_per_ray": round(b_ray, 1), "V": round(V, 3), "T": round(T, 3),
            "logical_gbs": round(logical_gbs, 1),
            "logical_frac": round(logical_gbs / HBM_PEAK_GBS, 4),
            "note": ("28 + 16 + 64 V + 36 T bytes per ray counts every node and triangle "
                     "fetch; caches serve most of them, so this is not an HBM rate"),
        },
    }
    if limiter:
        out["limiter"] = dict(limiter)
        out["limiter"]["statement"] = (
            "dependent node-fetch latency and divergence, not HBM bandwidth: a ray's "
            "next fetch needs the previous box test; lanes_per_valu of 64 shows the "
            "divergence, wait_share the share of wave cycles spent waiting")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; the traversal shards by rays "
                                   "(one replica of the tree per GPU) but no multi-GPU run "
                                   "is provided: use --gpus 1"}))
        sys.exit(2)

    import rta_amd
    from rta_amd import scenes
    from rta_amd.abi import NODE_DTYPE, TraceConfig, call, ptr
    from rta_amd.device import DeviceBuffer, Event, Stream

    if rta_amd.device_count() < 1:
        raise SystemExit("no HIP device visible")
    call("rta_hip_set_device", 0)
    dev_name = rta_amd.device_name()
    log(f"device {dev_name}")

    if args.workload == "dynamic":
        return dynamic_main(args)
    if args.workload == "batch":
        return batch_main(args)

    if args.copy_calibration:
        # a plain copy of known size: 1 GiB read + 1 GiB written per dispatch
        nbytes = 1 << 30
        src_b, dst_b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        st = Stream()
        a, b = Event(), Event()
        call("rta_hip_copy_kernel", ptr(dst_b), ptr(src_b), C.c_size_t(nbytes), ptr(st))
        times = []
        for _ in range(3):
            a.record(st)
            call("rta_hip_copy_kernel", ptr(dst_b), ptr(src_b), C.c_size_t(nbytes), ptr(st))
            b.record(st)
            b.synchronize()
            times.append(a.elapsed_ms(b))
        ms = min(times)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "ms": ms, "GB_s": 2 * nbytes / ms / 1e6}), flush=True)
        return

    t0 = time.perf_counter()
    tris_h = scenes.terrain_scene(grid=args.grid, n_random=args.random)
    n = len(tris_h)
    cam = scenes.terrain_camera(args.width, args.height)
    R = cam.w * cam.h
    log(f"scene {n} triangles, {R} primary rays ({time.perf_counter() - t0:.1f} s)")

    s = Stream()
    sp = ptr(s)
    pristine = DeviceBuffer.from_array(tris_h)
    tris = DeviceBuffer(104 * n)
    nodes = DeviceBuffer(32 * (2 * n - 1))
    idx = DeviceBuffer(4 * n)
    status = DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", n, C.byref(sz))
    ws = DeviceBuffer(sz.value)
    ws_bytes = C.c_size_t(sz.value)
    pairs = DeviceBuffer(64 * (2 * n - 1))
    tri4 = DeviceBuffer(48 * n)
    pflags = DeviceBuffer(4)
    prim = rta_amd.RayBuffers(R)
    bnc = rta_amd.RayBuffers(R)
    hits_p = DeviceBuffer(16 * R)
    hits_b = DeviceBuffer(16 * R)
    src = DeviceBuffer(4 * R)
    count = DeviceBuffer(4)
    call("rta_bounce_work_bytes", R, C.byref(sz))
    bwork = DeviceBuffer(sz.value)
    bwork_bytes = C.c_size_t(sz.value)
    err = DeviceBuffer(4)
    err.fill_u32(0)
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    vd, pos, up = f3(cam.view_dir), f3(cam.pos), f3(cam.up)
    fmax = float(rta_amd.FLT_MAX)

    # the root's link words are fixed by the scene (always an inner node
    # here); read them once after a first build
    # the build writes the traversal layout in its own passes
    # (rta_lbvh_build_packed), so build_ms covers everything the traversal
    # needs; the unfused pair (rta_lbvh_build, then rta_bvh_pack) is timed
    # once beside it
    def build_and_pack():
        call("rta_hip_memcpy_dtod", ptr(tris), ptr(pristine), C.c_size_t(104 * n), sp)
        ev[0].record(s)
        call("rta_lbvh_build_packed", ptr(tris), n, ptr(nodes), ptr(idx), ptr(pairs), ptr(tri4),
             ptr(pflags), ptr(ws), ws_bytes, ptr(status), None, sp)
        ev[1].record(s)
        ev[2].record(s)

    queue = DeviceBuffer(4)
    tcfg = TraceConfig(args.block, args.refill, args.steps_per_refill, 0, args.leaf_min)

    # the coherent primary rays gain from a short stack's extra resident
    # waves; the incoherent bounce rays lose L2 hits with them
    pcfg = TraceConfig(args.block, args.refill, args.steps_per_refill, 0, args.leaf_min,
                       args.primary_stack)
    povf = DeviceBuffer(4 * (R + 1))

    def trace(rays, ray_count, hits, root, primary=False):
        if args.trav == "packed":
            call("rta_trace_cis_packed", ptr(tri4), ptr(pairs), ptr(pflags), root[0], root[1],
                 ptr(rays.gpu_origin), ptr(rays.gpu_direction), ptr(rays.gpu_maxt), R,
                 ptr(ray_count), ptr(hits), ptr(err), sp)
        elif primary:
            call("rta_trace_cis_persistent_ovf", ptr(tri4), ptr(pairs), ptr(pflags), root[0],
                 root[1], ptr(rays.gpu_origin), ptr(rays.gpu_direction), ptr(rays.gpu_maxt), R,
                 ptr(ray_count), ptr(hits), ptr(err), ptr(queue), C.byref(pcfg), ptr(povf), sp)
        else:
            call("rta_trace_cis_persistent", ptr(tri4), ptr(pairs), ptr(pflags), root[0], root[1],
                 ptr(rays.gpu_origin), ptr(rays.gpu_direction), ptr(rays.gpu_maxt), R,
                 ptr(ray_count), ptr(hits), ptr(err), ptr(queue), C.byref(tcfg), sp)

    def frame(root):
        build_and_pack()
        call("rta_setup_shirley_rays", ptr(prim.gpu_direction), ptr(prim.gpu_origin),
             ptr(prim.gpu_maxt), cam.fovy, cam.aspect, cam.w, cam.h, vd, pos, up, fmax, sp)
        call("rta_reset_intersections", ptr(hits_p), R, 1, sp)
        ev[3].record(s)
        trace(prim, None, hits_p, root, primary=True)
        ev[4].record(s)
        call("rta_diffuse_bounce", ptr(bnc.gpu_origin), ptr(bnc.gpu_direction), ptr(bnc.gpu_maxt),
             ptr(src), ptr(count), ptr(hits_p), ptr(tris), ptr(prim.gpu_origin),
             ptr(prim.gpu_direction), R, C.c_uint64(scenes.BOUNCE_SEED), ptr(bwork), bwork_bytes,
             sp)
        call("rta_reset_intersections", ptr(hits_b), R, 1, sp)
        ev[5].record(s)
        trace(bnc, count, hits_b, root)
        ev[6].record(s)

    ev = [Event() for _ in range(7)]
    build_and_pack()
    s.synchronize()
    r0 = nodes.download(NODE_DTYPE, 1)[0]
    root = (int(r0["type_left_elems"]), int(r0["right_tris"]))

    log(f"warmup {args.warmup} steps")
    for _ in range(args.warmup):
        frame(root)
    s.synchronize()

    seg = {"build": [], "primary": [], "bounce_gen": [], "bounce": []}
    log(f"timing {args.steps} steps")
    call("rta_hip_device_synchronize")
    t_start = time.perf_counter()
    for _ in range(args.steps):
        frame(root)
        ev[6].synchronize()
        seg["build"].append(ev[0].elapsed_ms(ev[1]))
        seg["primary"].append(ev[3].elapsed_ms(ev[4]))
        seg["bounce_gen"].append(ev[4].elapsed_ms(ev[5]))
        seg["bounce"].append(ev[5].elapsed_ms(ev[6]))
    call("rta_hip_device_synchronize")
    wall_ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    K = args.steps
    avg = {k: float(np.mean(v)) for k, v in seg.items()}
    med = {k: float(np.median(v)) for k, v in seg.items()}

    # the reference's own convention for trace_rays (bbvh-cuda.h:221-229):
    # host wall clock around launch + synchronisation, reset untimed
    walls = []
    for _ in range(max(K, 10)):
        call("rta_reset_intersections", ptr(hits_b), R, 1, sp)
        s.synchronize()
        t0 = time.perf_counter()
        trace(bnc, count, hits_b, root)
        s.synchronize()
        walls.append((time.perf_counter() - t0) * 1e3)
    bounce_wall_ms = float(np.median(walls))

    e = int(err.download(np.uint32, 1)[0])
    if e:
        raise SystemExit(f"traversal reported {e} stack/guard errors")
    n_bounce = int(count.download(np.uint32, 1)[0])
    st = int(status.download(np.uint32, 1)[0])

    # per-stage build breakdown (one extra, synchronising build)
    times = rta_amd.abi.RtaTimes()
    call("rta_hip_memcpy_dtod", ptr(tris), ptr(pristine), C.c_size_t(104 * n), sp)
    call("rta_lbvh_build_packed", ptr(tris), n, ptr(nodes), ptr(idx), ptr(pairs), ptr(tri4),
         ptr(pflags), ptr(ws), ws_bytes, ptr(status), C.byref(times), sp)
    stage_ms = times.as_dict()
    # the unfused pair: the seven stages, then the layout derived from the
    # finished nodes (medians of 10)
    unfused = {"build": [], "pack": []}
    for _ in range(10):
        call("rta_hip_memcpy_dtod", ptr(tris), ptr(pristine), C.c_size_t(104 * n), sp)
        ev[0].record(s)
        call("rta_lbvh_build", ptr(tris), n, ptr(nodes), ptr(idx), ptr(ws), ws_bytes,
             ptr(status), None, sp)
        ev[1].record(s)
        call("rta_bvh_pack", ptr(pairs), ptr(tri4), ptr(pflags), ptr(nodes), 2 * n - 1,
             ptr(tris), n, sp)
        ev[2].record(s)
        ev[2].synchronize()
        unfused["build"].append(ev[0].elapsed_ms(ev[1]))
        unfused["pack"].append(ev[1].elapsed_ms(ev[2]))
    unfused = {k: float(np.median(v)) for k, v in unfused.items()}

    # per-ray inner-node visits V and triangle tests T on the bounce rays,
    # from the instrumented plain kernel (equal to the oracle's counts, see
    # tests/test_gpu_parity.py::test_trace_sphere_bit_exact)
    stats = DeviceBuffer(8 * max(n_bounce, 1))
    call("rta_reset_intersections", ptr(hits_b), R, 1, sp)
    call("rta_trace_cis", ptr(tris), n, ptr(nodes), ptr(bnc.gpu_origin), ptr(bnc.gpu_direction),
         ptr(bnc.gpu_maxt), n_bounce, 1, ptr(hits_b), ptr(stats), ptr(err), sp)
    s.synchronize()
    vt = stats.download(np.uint32, 2 * n_bounce).reshape(-1, 2).astype(np.float64)
    V, T = float(vt[:, 0].mean()), float(vt[:, 1].mean())
    b_ray = 28 + 16 + 64 * V + 36 * T
    achieved = b_ray * n_bounce / (avg["bounce"] * 1e-3) / 1e9

    bounce_mrays = n_bounce / (avg["bounce"] * 1e3)
    primary_mrays = R / (avg["primary"] * 1e3)
    build_mtri = n / (avg["build"] * 1e3)
    log(f"primary {primary_mrays:.1f} Mrays/s, bounce {bounce_mrays:.1f} Mrays/s, "
        f"build {build_mtri:.1f} Mtri/s ({avg['build']:.3f} ms), V={V:.1f} T={T:.2f}")

    # HBM traffic per bounce launch from the committed rocprofv3 --pmc
    # measurement of this same workload (FETCH_SIZE and WRITE_SIZE, each
    # scaled by what the counter reads for a copy kernel of known size)
    traffic, traffic_src, limiter = None, None, None
    pmc_file = os.path.join(ROOT, "profiles", "trace_persistent_heightfield_2M_pmc.json")
    default_workload = (args.trav == "persistent" and args.grid == 1024 and
                        args.random == 200_000 and (args.width, args.height) == (1920, 1080))
    if default_workload and os.path.exists(pmc_file):
        with open(pmc_file) as f:
            pmc = json.load(f)
        rec = pmc["dispatches"].get("bounce rays")
        if rec:
            traffic = round(rec["fetch_bytes"] + rec["write_bytes"])
            traffic_src = os.path.relpath(pmc_file, ROOT)
            limiter = {k: rec[k] for k in ("l2_hit_pct", "valu_issue_share", "lanes_per_valu",
                                           "wait_share") if k in rec}

    # the build's HBM traffic per launch from the committed rocprofv3 --pmc
    # measurement of this workload's packed build (same scaling)
    build_traffic = None
    bpmc_file = os.path.join(ROOT, "profiles", "lbvh_build_heightfield_2M_pmc.json")
    if default_workload and os.path.exists(bpmc_file):
        with open(bpmc_file) as f:
            build_traffic = int(json.load(f)["total_bytes"])

    hit_predicted = None
    if not args.no_predictor and default_workload:
        from rta_amd.bvh import PackedBvh

        packed = PackedBvh(pairs, tri4, pflags, root[0], root[1])
        ranges = DeviceBuffer(8 * (n - 1))
        call("rta_lbvh_copy_ranges", ptr(ranges), ptr(ws), n, sp)
        hit_predicted = predictor_leg(args, tris_h, tris, nodes, packed, ranges, prim, hits_p, R,
                                      tcfg, s)
        del ranges

    # the many-mesh and dynamic-frame workloads, timed in the same run
    batch = dynamic = None
    if default_workload and not args.no_legs:
        batch = batch_leg(args)
        dynamic = dynamic_leg(args)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle restatement of the reference CPU plugin)")
        o, d, t = bnc.download(n_bounce)
        cpu = cpu_baseline(tris_h, o, d, t, n_bounce, cam.w, args.cpu_rays)
        log(f"CPU {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(bounce_mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("2,297,152-triangle heightfield+random scene, LBVH build + 1920x1080 "
                         "primary rays + one diffuse bounce per hit; value = plain closest-hit "
                         "Mrays/s on the bounce rays"),
            "n_tris": n, "primary_rays": R, "bounce_rays": n_bounce,
            "traversal": (f"child-is, packed 64-B child-pair nodes, LDS stack, {args.trav}"
                          + (f" (block {args.block}, refill {args.refill}, "
                             f"steps {args.steps_per_refill}, leaf_min {args.leaf_min}; "
                             f"stack 32 for the bounce rays, {args.primary_stack} for the "
                             f"primary rays)"
                             if args.trav == "persistent" else "")),
        },
        "primary_mrays_s": round(primary_mrays, 3),
        "bounce_mrays_s": round(bounce_mrays, 3),
        "build_mtri_s": round(build_mtri, 3),
        # compulsory bytes of the build (read 104 B triangles, write 32 B
        # nodes, 4 B indices and the 104 B reordered triangles) over its time
        "build_roofline": {
            "bound": "hbm", "bytes_per_tri": 276, "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved": round(276 * n / (avg["build"] * 1e-3) / 1e9, 2),
            "frac": round(276 * n / (avg["build"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            # measured bytes between L2 and the fabric per build (all stages)
            "traffic": build_traffic,
            "traffic_frac": (round(build_traffic / (avg["build"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                             if build_traffic else None),
        },
        "build_ms": round(avg["build"], 4),
        "build_includes_traversal_layout": True,
        "unfused_build_ms": round(unfused["build"], 4),
        "unfused_pack_ms": round(unfused["pack"], 4),
        "bounce_gen_ms": round(avg["bounce_gen"], 4),
        "primary_trace_ms": round(avg["primary"], 4),
        "bounce_trace_ms": round(avg["bounce"], 4),
        "median_ms": {k: round(v, 4) for k, v in med.items()},
        "bounce_mrays_s_median": round(n_bounce / (med["bounce"] * 1e3), 3),
        # reference-style trace time: launch + kernel + synchronisation
        "bounce_trace_wall_ms": round(bounce_wall_ms, 4),
        "bounce_mrays_s_wall": round(n_bounce / (bounce_wall_ms * 1e3), 3),
        "build_stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "lbvh_status": st,
        "roofline": traversal_roofline(args, traffic, traffic_src, limiter, avg["bounce"], b_ray,
                                       achieved, V, T),
        "hit_predicted": hit_predicted,
        "batch": batch,
        "dynamic": dynamic,
        "cpu_baseline": cpu,
        "device": dev_name,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
