"""GPU parity: every stage of the HIP LBVH build, the traversal kernels, the
ray generator and the bouncer against the CPU oracle, bit for bit.

Bar: integer/index outputs (codes, permutation, links, parents, hit ids)
identical; float outputs (boxes, centroids, t, beta, gamma, rays) identical
as IEEE bit patterns (a 1e-6 relative tolerance on t and the barycentrics
would be acceptable; these kernels meet it with zero difference), except that the
sign of a zero inside a box is compared by value (-0 == +0, which the
reference leaves to thread arrival order)."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle
import rta_amd
from rta_amd import scenes
from rta_amd.abi import AABB_DTYPE, HIT_DTYPE, NODE_DTYPE, TRI_DTYPE, TraceConfig, call, ptr
from rta_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
F32MAX = np.float32(np.finfo(np.float32).max)


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def clustered(n, cells, seed):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, TRI_DTYPE)
    centres = rng.uniform(0, 1, size=(cells, 3)).astype(np.float32)
    base = centres[rng.integers(0, cells, size=n)]
    for k in ("a", "b", "c"):
        t[k] = base + rng.uniform(-1e-5, 1e-5, size=(n, 3)).astype(np.float32)
    t[0]["a"], t[0]["b"], t[0]["c"] = (0, 0, 0), (1e-4, 0, 0), (0, 1e-4, 1e-4)
    t[1]["a"], t[1]["b"], t[1]["c"] = (1, 1, 1), (0.99, 1, 1), (1, 0.99, 0.99)
    return t


SCENES = {
    "icosphere2": lambda: scenes.icosphere(2),
    "clustered": lambda: clustered(5000, 37, seed=5),
    "random_small": lambda: scenes.random_triangles(777, seed=3),
    "sphere20k": scenes.sphere_scene,
    "hf_small": lambda: np.concatenate([scenes.heightfield(64, 64, seed=3),
                                        scenes.random_triangles(3000, seed=4)]),
}


def gpu_stages(tris):
    """Run the seven stages one entry point at a time; return intermediates."""
    n = len(tris)
    src = DeviceBuffer.from_array(tris)
    scratch = DeviceBuffer(104 * n)
    boxes = DeviceBuffer(24 * n)
    centers = DeviceBuffer(12 * n)
    partials = DeviceBuffer(24 * ((n + 255) // 256))
    bounds = DeviceBuffer(24)
    codes = DeviceBuffer(4 * n)
    idx = DeviceBuffer(4 * n)
    status = DeviceBuffer(4)
    status.fill_u32(0)
    nodes = DeviceBuffer(32 * (2 * n - 1))
    parents = DeviceBuffer(4 * (2 * n - 1))
    sz = C.c_size_t(0)
    call("rta_lbvh_boxes_work_bytes", n, C.byref(sz))
    visit = DeviceBuffer(sz.value)
    call("rta_lbvh_sort_work_bytes", n, C.byref(sz))
    work = DeviceBuffer(sz.value)
    out = {}
    call("rta_lbvh_init_triangles", ptr(scratch), ptr(src), ptr(boxes), ptr(centers),
         ptr(partials), n, None)
    out["scratch"] = scratch.download(TRI_DTYPE, n)
    out["boxes"] = boxes.download(AABB_DTYPE, n)
    out["centers"] = centers.download(np.float32, 3 * n)
    call("rta_lbvh_voxel_bounds", ptr(bounds), ptr(partials), n, None)
    out["bounds"] = bounds.download(AABB_DTYPE, 1)
    call("rta_lbvh_morton_codes", ptr(codes), ptr(idx), ptr(centers), ptr(bounds), ptr(status), n,
         None)
    out["codes_input"] = codes.download(np.uint32, n)
    out["status"] = int(status.download(np.uint32, 1)[0])
    call("rta_lbvh_sort", ptr(codes), ptr(idx), n, ptr(work), C.c_size_t(sz.value), None)
    out["codes_sorted"] = codes.download(np.uint32, n)
    out["indices"] = idx.download(np.uint32, n)
    call("rta_lbvh_build_tree", ptr(nodes), ptr(parents), ptr(visit), ptr(codes), n, None)
    out["links"] = nodes.download(NODE_DTYPE, 2 * n - 1)
    out["parents"] = parents.download(np.uint32, 2 * n - 1)
    call("rta_lbvh_build_boxes", ptr(nodes), ptr(visit), ptr(boxes), ptr(codes), ptr(idx),
         ptr(parents), n, None)
    out["nodes"] = nodes.download(NODE_DTYPE, 2 * n - 1)
    call("rta_lbvh_permute", ptr(src), ptr(scratch), ptr(idx), n, None)
    out["tris"] = src.download(TRI_DTYPE, n)
    return out


def assert_nodes_equal(got, exp):
    assert np.array_equal(got["type_left_elems"], exp["type_left_elems"])
    assert np.array_equal(got["right_tris"], exp["right_tris"])
    assert np.array_equal(got["min"], exp["min"])
    assert np.array_equal(got["max"], exp["max"])


@pytest.mark.parametrize("name", list(SCENES))
def test_lbvh_stages_bit_exact(gpu, name):
    tris = SCENES[name]()
    n = len(tris)
    say(f"lbvh stages scene={name} n={n}")
    ref = oracle.lbvh_build(tris)
    got = gpu_stages(tris)
    assert np.array_equal(bits(got["scratch"]), bits(tris))
    v = np.stack([tris["a"], tris["b"], tris["c"]], axis=1)
    assert np.array_equal(got["boxes"]["min"], v.min(axis=1))
    assert np.array_equal(got["boxes"]["max"], v.max(axis=1))
    ctr = ((tris["a"] + tris["b"]) + tris["c"]) / np.float32(3)
    assert np.array_equal(bits(got["centers"]), bits(ctr.reshape(-1)))
    assert np.array_equal(got["bounds"]["min"], ref["bounds"]["min"])
    assert np.array_equal(got["bounds"]["max"], ref["bounds"]["max"])
    assert np.array_equal(got["codes_input"], ref["codes_input"])
    assert got["status"] == ref["status"]
    assert np.array_equal(got["codes_sorted"], ref["codes_sorted"])
    assert np.array_equal(got["indices"], ref["indices"])
    assert np.array_equal(got["parents"][1:], ref["parents"][1:])
    assert_nodes_equal(got["nodes"], ref["nodes"])
    assert np.array_equal(bits(got["tris"]), bits(ref["tris"]))


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 4095, 4096, 4097, 100_003])
def test_lbvh_full_build_sizes(gpu, n):
    say(f"lbvh full build n={n}")
    tris = scenes.random_triangles(n, seed=n)
    ref = oracle.lbvh_build(tris)
    dev = DeviceBuffer.from_array(tris)
    ctor = rta_amd.LbvhConstructor()
    bvh = ctor.build(dev, n)
    assert ctor.last_status == ref["status"]
    assert_nodes_equal(bvh.nodes(), ref["nodes"])
    assert np.array_equal(bvh.indices(), ref["indices"])
    assert np.array_equal(bits(bvh.canonical_triangle_ptr()), bits(ref["tris"]))
    assert set(ctor.last_times) >= {"sort", "tree", "boxes", "total"}


def test_lbvh_rejects_empty(gpu):
    ws = DeviceBuffer(1024)
    with pytest.raises(rta_amd.RtaError):
        call("rta_lbvh_build", ptr(ws), 0, ptr(ws), ptr(ws), ptr(ws), C.c_size_t(1024), None,
             None, None)
    tris = DeviceBuffer.from_array(scenes.random_triangles(1000, seed=1))
    with pytest.raises(rta_amd.RtaError) as e:
        call("rta_lbvh_build", ptr(tris), 1000, ptr(ws), ptr(ws), ptr(ws), C.c_size_t(1024), None,
             None, None)
    assert e.value.code == -2


def test_lbvh_rebuild_same_buffers(gpu):
    """lbvh.h:157-162: a same-n rebuild reuses the buffers; results stay exact."""
    ctor = rta_amd.LbvhConstructor()
    for seed in (1, 2):
        say(f"rebuild seed={seed}")
        tris = scenes.icosphere(4, jitter=0.02, seed=seed)
        ref = oracle.lbvh_build(tris)
        bvh = ctor.build(DeviceBuffer.from_array(tris), len(tris))
        assert_nodes_equal(bvh.nodes(), ref["nodes"])


@pytest.mark.parametrize("n,dup", [(1, 1), (5, 2), (4096, 7), (4097, 100), (1_000_000, 50_000),
                                   (3_000_001, 1 << 30)])
def test_radix_sort_stable(gpu, n, dup):
    say(f"radix sort n={n} dup={dup}")
    rng = np.random.default_rng(n)
    keys = rng.integers(0, dup, size=n, dtype=np.int64).astype(np.uint32) & np.uint32(0x3FFFFFFF)
    vals = np.arange(n, dtype=np.uint32)
    kd, vd = DeviceBuffer.from_array(keys), DeviceBuffer.from_array(vals)
    sz = C.c_size_t(0)
    call("rta_lbvh_sort_work_bytes", n, C.byref(sz))
    work = DeviceBuffer(sz.value)
    call("rta_lbvh_sort", ptr(kd), ptr(vd), n, ptr(work), C.c_size_t(sz.value), None)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(kd.download(np.uint32, n), keys[order])
    assert np.array_equal(vd.download(np.uint32, n), order.astype(np.uint32))


def test_shirley_rays_match_host(gpu):
    cam = scenes.terrain_camera(333, 177)
    o, d, t = cam.host_rays()
    rg = rta_amd.ShirleyRayGenerator(cam.w, cam.h)
    rg.setup(cam.pos, cam.view_dir, cam.up, cam.fovy)
    rg.generate_rays()
    go, gd, gt = rg.download()
    assert np.array_equal(bits(go), bits(o))
    assert np.array_equal(bits(gd), bits(d))
    assert np.array_equal(bits(gt), bits(t))


def test_reset_sentinel(gpu):
    b = rta_amd.GpuBouncer(1000)
    b.gpu_last_intersection.fill_u32(0xDEADBEEF)
    call("rta_reset_intersections", ptr(b.gpu_last_intersection), 10, 100, None)
    h = b.intersections()
    assert (h["t"] == F32MAX).all() and (h["beta"] == 1).all() and (h["gamma"] == 2).all()
    assert (h["ref"] == 0xDEADBEEF).all()   # ref untouched (cuda-kernels.cu:26-28)


LAYOUTS = ["plain", "packed", "persistent"]


def trace_gpu(bvh, o, d, t, layout, stats=False, config=None):
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    bouncer = rta_amd.GpuBouncer(len(t))
    tracer = rta_amd.CisTracer(rays, bvh, bouncer, layout=layout, config=config)
    tracer.prepare_trace()
    st = DeviceBuffer(8 * len(t)) if stats else None
    tracer.launch(stats=st)
    rta_amd.synchronize()
    tracer.check_errors()
    hits = bouncer.intersections()
    return hits, (st.download(np.uint32, 2 * len(t)) if stats else None)


def assert_hits_equal(got, ref):
    hit = ref["t"] != F32MAX
    assert np.array_equal(got["t"] != F32MAX, hit)
    assert np.array_equal(bits(got["t"][hit]), bits(ref["t"][hit]))
    assert np.array_equal(bits(got["beta"][hit]), bits(ref["beta"][hit]))
    assert np.array_equal(bits(got["gamma"][hit]), bits(ref["gamma"][hit]))
    assert np.array_equal(got["ref"][hit], ref["ref"][hit])
    # a miss keeps the reset sentinel (kernels.cu:793-794, cuda-kernels.cu:26-28)
    assert (got["beta"][~hit] == 1).all() and (got["gamma"][~hit] == 2).all()


@pytest.mark.parametrize("layout", LAYOUTS)
def test_trace_sphere20k_bit_exact(gpu, layout):
    say(f"trace sphere20k layout={layout}")
    tris = scenes.sphere_scene()
    ref_b = oracle.lbvh_build(tris)
    cam = scenes.sphere_camera()
    o, d, t = cam.host_rays()
    ref, _, ref_st = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], o, d, t, cam.w, cam.h,
                                      stats=True)
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    got, st = trace_gpu(bvh, o, d, t, layout, stats=(layout == "plain"))
    assert (ref["t"] != F32MAX).mean() > 0.2
    assert_hits_equal(got, ref)
    if st is not None:
        assert np.array_equal(st, ref_st)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_trace_incoherent_and_duplicates(gpu, layout):
    say(f"trace incoherent layout={layout}")
    tris = np.concatenate([clustered(4000, 23, seed=8), scenes.icosphere(3, jitter=0.05, seed=2)])
    ref_b = oracle.lbvh_build(tris)
    rng = np.random.default_rng(4)
    n = 50_000
    o = rng.uniform(-1.5, 1.5, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[: n // 10, 0] = 0.0                     # axis-parallel rays: d == 0 slabs
    d[n // 10::7, 1] = 0.0                    # and single ones inside waves of others
    d[n // 10::13, 2] = 0.0
    o[n // 10::11, 0] = np.float32(1e-30)     # origins outside the exact-division range
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    o, d = o.reshape(-1), d.reshape(-1)
    t = np.full(n, F32MAX, np.float32)
    t[: n // 7] = np.float32(0.75)            # finite max_t for some rays
    ref, _, _ = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], o, d, t, n)
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    got, _ = trace_gpu(bvh, o, d, t, layout)
    # the GPU kernel starts closest.t at FLT_MAX (kernels.cu:794), the CPU
    # tracer at max_t (bbvh.h:412): identical hits, a finite-max_t miss
    # reports FLT_MAX on the GPU and max_t on the CPU
    ref = ref.copy()
    miss_cpu = ref["t"] == t
    ref["t"][miss_cpu] = F32MAX
    assert_hits_equal(got, ref)


def test_trace_median_tree(gpu):
    """The oracle's own spatial-median tree traced on the GPU (a parity check
    independent of the LBVH, bbvh-cuda/main.cpp:158-163)."""
    say("trace median tree")
    tris = scenes.icosphere(4, jitter=0.03, seed=9)
    mb = oracle.median_build(tris)
    cam = scenes.sphere_camera(200, 200)
    o, d, t = cam.host_rays()
    ref, _, _ = oracle.trace_cis(mb["tris"], mb["nodes"], o, d, t, 200, 200)
    bvh = rta_amd.BinaryLbvh(DeviceBuffer.from_array(mb["tris"]), len(tris),
                             DeviceBuffer.from_array(mb["nodes"]), DeviceBuffer(4 * len(tris)))
    # a median tree has fewer nodes than 2n-1: pad the node array
    nodes = np.zeros(2 * len(tris) - 1, NODE_DTYPE)
    nodes[: len(mb["nodes"])] = mb["nodes"]
    bvh.node_data = DeviceBuffer.from_array(nodes)
    for layout in LAYOUTS:
        got, _ = trace_gpu(bvh, o, d, t, layout)
        assert_hits_equal(got, ref)
        bvh.invalidate()


@pytest.mark.parametrize("shape", [(64, 1, 1, 0, 1), (64, 64, 1, 0, 64), (128, 16, 7, 0, 3),
                                   (256, 33, 100, 0, 40), (64, 8, 32, 0, 0), (64, 8, 32, 1, 16)])
def test_persistent_launch_shapes(gpu, shape):
    """Every launch shape of the persistent tracer gives the oracle's hits."""
    say(f"persistent shape={shape}")
    tris = scenes.icosphere(4, jitter=0.05, seed=3)
    ref_b = oracle.lbvh_build(tris)
    rng = np.random.default_rng(9)
    n = 30_001
    o = rng.uniform(-1.2, 1.2, size=(n, 3)).astype(np.float32).reshape(-1)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32).reshape(-1)
    t = np.full(n, F32MAX, np.float32)
    ref, _, _ = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], o, d, t, n)
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    got, _ = trace_gpu(bvh, o, d, t, "persistent", config=TraceConfig(*shape))
    assert_hits_equal(got, ref)


def test_bounce_bit_exact(gpu):
    say("bounce")
    tris = scenes.sphere_scene()
    ref_b = oracle.lbvh_build(tris)
    cam = scenes.sphere_camera(256, 256)
    o, d, t = cam.host_rays()
    ref, _, _ = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], o, d, t, 256, 256)
    bo, bd, bt, bsrc = oracle.diffuse_bounce(ref, ref_b["tris"], o, d, seed=scenes.BOUNCE_SEED)
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    bouncer = rta_amd.GpuBouncer(len(t))
    tracer = rta_amd.CisTracer(rays, bvh, bouncer)
    tracer.prepare_trace()
    tracer.launch()
    out, src, count = rta_amd.diffuse_bounce(bouncer.gpu_last_intersection, bvh.triangle_data,
                                             rays, len(t), scenes.BOUNCE_SEED)
    assert count == len(bt)
    go, gd, gt = out.download(count)
    assert np.array_equal(src.download(np.uint32, count), bsrc)
    assert np.array_equal(bits(go), bits(bo))
    assert np.array_equal(bits(gd), bits(bd))
    assert np.array_equal(bits(gt), bits(bt))
    # and the bounce rays trace identically
    ref2, _, _ = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], bo, bd, bt, count)
    got2, _ = trace_gpu(bvh, go, gd, gt, "packed")
    assert_hits_equal(got2, ref2)


# ------------------------------- shadow tracer, finite max_t, ray chains

def _mixed_scene():
    return np.concatenate([scenes.icosphere(3, jitter=0.05, seed=21),
                           scenes.random_triangles(4000, seed=22, lo=-1.2, hi=1.2)])


def _segment_rays(n, seed):
    """Rays between random point pairs: finite max_t = segment length, plus
    every fifth ray unbounded."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-1.5, 1.5, size=(n, 3)).astype(np.float32)
    p = rng.uniform(-1.5, 1.5, size=(n, 3)).astype(np.float32)
    v = (p - o).astype(np.float32)
    ln = np.linalg.norm(v, axis=1).astype(np.float32)
    d = (v / ln[:, None]).astype(np.float32)
    ln[::5] = F32MAX
    d[n // 2::9, 0] = 0.0                       # some axis-parallel rays
    return o.reshape(-1), d.reshape(-1), ln


def _trace_gpu_kind(bvh, o, d, t, layout, shadow):
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    bouncer = rta_amd.GpuBouncer(len(t))
    cls = rta_amd.CisShadowTracer if shadow else rta_amd.CisTracer
    tracer = cls(rays, bvh, bouncer, layout=layout)
    tracer.prepare_trace()
    tracer.launch()
    rta_amd.synchronize()
    tracer.check_errors()
    return bouncer.intersections()


def _assert_records_equal(got, exp):
    """t, beta, gamma of every record; ref where the record is a hit (a miss
    keeps the stale ref the buffer held, cuda-kernels.cu:26-28)."""
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(got[f]), bits(exp[f])), f
    hit = exp["t"] != F32MAX
    assert np.array_equal(got["ref"][hit], exp["ref"][hit])


@pytest.mark.parametrize("layout", ["packed", "persistent"])
@pytest.mark.parametrize("shadow", [False, True])
def test_kernel_semantics_finite_max_t(gpu, layout, shadow):
    """Closest-hit and shadow traversal against the oracle's restatement of
    the GPU kernels, every field of every record (misses keep the reset
    sentinel apart from t)."""
    say(f"kernel semantics layout={layout} shadow={shadow}")
    tris = _mixed_scene()
    ref_b = oracle.lbvh_build(tris)
    o, d, t = _segment_rays(40_000, seed=23)
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    got = _trace_gpu_kind(bvh, o, d, t, layout, shadow)
    exp, deep = oracle.trace_kernel(ref_b["tris"], ref_b["nodes"], o, d, t, any_hit=shadow)
    assert deep == 0
    hit = exp["t"] != F32MAX
    assert 5000 < hit.sum() < len(t) - 5000
    # every record: t, beta, gamma (a miss is the reset sentinel t = FLT_MAX,
    # beta = 1, gamma = 2); ref on hits (a miss keeps the buffer's stale ref,
    # cuda-kernels.cu:26-28)
    for f in ("t", "beta", "gamma"):
        assert np.array_equal(bits(got[f]), bits(exp[f])), f
    assert np.array_equal(got["ref"][hit], exp["ref"][hit])


def test_shadow_rays_bit_exact(gpu):
    say("shadow rays")
    tris = scenes.sphere_scene()
    ref_b = oracle.lbvh_build(tris)
    cam = scenes.sphere_camera(256, 256)
    o, d, t = cam.host_rays()
    ref, _, _ = oracle.trace_cis(ref_b["tris"], ref_b["nodes"], o, d, t, 256, 256)
    light = (4.0, 0.0, 1.0)
    so, sd, st, ssrc = oracle.shadow_rays(ref, ref_b["tris"], o, d, np.float32(light))
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    bouncer = rta_amd.GpuBouncer(len(t))
    tracer = rta_amd.CisTracer(rays, bvh, bouncer)
    tracer.prepare_trace()
    tracer.launch()
    out, src, count = rta_amd.shadow_rays(bouncer.gpu_last_intersection, bvh.triangle_data, rays,
                                          len(t), light)
    assert count == len(st)
    go, gd, gt = out.download(count)
    assert np.array_equal(src.download(np.uint32, count), ssrc)
    assert np.array_equal(bits(go), bits(so))
    assert np.array_equal(bits(gd), bits(sd))
    assert np.array_equal(bits(gt), bits(st))
    exp, _ = oracle.trace_kernel(ref_b["tris"], ref_b["nodes"], so, sd, st, any_hit=True)
    got = _trace_gpu_kind(bvh, go, gd, gt, "persistent", shadow=True)
    hit = exp["t"] != F32MAX
    assert 1000 < hit.sum() < count - 1000
    _assert_records_equal(got[hit], exp[hit])
    assert (got["t"][~hit] == F32MAX).all()


def test_two_bounce_chain(gpu):
    """primary -> bounce -> trace -> bounce -> trace entirely on the GPU
    (ray counts stay on the device) against the same chain in the oracle."""
    say("two-bounce chain")
    tris = np.concatenate([scenes.heightfield(96, 96, seed=31),
                           scenes.random_triangles(5000, seed=32)])
    ref_b = oracle.lbvh_build(tris)
    cam = scenes.Camera(pos=(0.5, -0.3, 0.6), target=(0.5, 0.5, 0.0), up=(0, 0, 1), fovy=50.0,
                        w=192, h=128)
    o, d, t = cam.host_rays()
    R = len(t)
    bvh = rta_amd.LbvhConstructor().build(DeviceBuffer.from_array(tris), len(tris))
    rays = rta_amd.RayBuffers.from_host(o, d, t)
    hits = [rta_amd.GpuBouncer(R) for _ in range(3)]
    tr = rta_amd.CisTracer(rays, bvh, hits[0])
    tr.prepare_trace()
    tr.launch()
    exp0, _ = oracle.trace_kernel(ref_b["tris"], ref_b["nodes"], o, d, t)
    cur_rays, cur_o, cur_d, cur_exp = rays, o, d, exp0
    for k in (1, 2):
        seed = scenes.BOUNCE_SEED + k
        b_rays, _, count = rta_amd.diffuse_bounce(hits[k - 1].gpu_last_intersection,
                                                  bvh.triangle_data, cur_rays, R, seed)
        bo, bd, bt, _ = oracle.diffuse_bounce(cur_exp, ref_b["tris"], cur_o, cur_d, seed)
        assert count == len(bt) and count > 1000
        go, gd, gt = b_rays.download(count)
        assert np.array_equal(bits(go), bits(bo)) and np.array_equal(bits(gd), bits(bd))
        tr = rta_amd.CisTracer(b_rays, bvh, hits[k])
        tr.prepare_trace()                      # resets all R records
        tr.launch(n_rays=count)
        rta_amd.synchronize()
        tr.check_errors()
        got = hits[k].intersections()
        exp, _ = oracle.trace_kernel(ref_b["tris"], ref_b["nodes"], bo, bd, bt)
        _assert_records_equal(got[:count], exp)
        assert (got["t"][count:] == F32MAX).all()
        cur_rays, cur_o, cur_d = b_rays, go, gd
        cur_exp = np.concatenate([exp, got[count:]])
        cur_exp["t"][count:] = F32MAX


def _build_then_pack(tris):
    n = len(tris)
    t = DeviceBuffer.from_array(tris)
    nodes, idx, st = DeviceBuffer(32 * (2 * n - 1)), DeviceBuffer(4 * n), DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", n, C.byref(sz))
    ws = DeviceBuffer(sz.value)
    call("rta_lbvh_build", ptr(t), n, ptr(nodes), ptr(idx), ptr(ws), C.c_size_t(sz.value),
         ptr(st), None, None)
    pairs, tri4, flags = DeviceBuffer(64 * (2 * n - 1)), DeviceBuffer(48 * n), DeviceBuffer(4)
    call("rta_bvh_pack", ptr(pairs), ptr(tri4), ptr(flags), ptr(nodes), 2 * n - 1, ptr(t), n,
         None)
    rta_amd.synchronize()
    return {"nodes": nodes.download(NODE_DTYPE, 2 * n - 1), "idx": idx.download(np.uint32, n),
            "tris": t.download(TRI_DTYPE, n), "status": int(st.download(np.uint32, 1)[0]),
            "pairs": pairs.download(np.uint8, 64 * (2 * n - 1)).reshape(-1, 64),
            "tri4": tri4.download(np.uint8, 48 * n), "flags": int(flags.download(np.uint32, 1)[0])}


def _build_packed(tris):
    n = len(tris)
    t = DeviceBuffer.from_array(tris)
    nodes, idx, st = DeviceBuffer(32 * (2 * n - 1)), DeviceBuffer(4 * n), DeviceBuffer(4)
    sz = C.c_size_t(0)
    call("rta_lbvh_workspace_bytes", n, C.byref(sz))
    ws = DeviceBuffer(sz.value)
    pairs, tri4, flags = DeviceBuffer(64 * (2 * n - 1)), DeviceBuffer(48 * n), DeviceBuffer(4)
    flags.fill_u32(0xffffffff)               # the build must clear it
    tri4.fill_u32(0x7fc00000)                # and write every tri4 word
    call("rta_lbvh_build_packed", ptr(t), n, ptr(nodes), ptr(idx), ptr(pairs), ptr(tri4),
         ptr(flags), ptr(ws), C.c_size_t(sz.value), ptr(st), None, None)
    rta_amd.synchronize()
    return {"nodes": nodes.download(NODE_DTYPE, 2 * n - 1), "idx": idx.download(np.uint32, n),
            "tris": t.download(TRI_DTYPE, n), "status": int(st.download(np.uint32, 1)[0]),
            "pairs": pairs.download(np.uint8, 64 * (2 * n - 1)).reshape(-1, 64),
            "tri4": tri4.download(np.uint8, 48 * n), "flags": int(flags.download(np.uint32, 1)[0])}


PACKED_SCENES = dict(SCENES, **{
    "one": lambda: scenes.random_triangles(1, seed=1),
    "two": lambda: scenes.random_triangles(2, seed=2),
    "dup_heavy": lambda: clustered(70_000, 5, seed=8),        # long collapsed runs
    "two_merge_levels": lambda: scenes.random_triangles(100_003, seed=12),
    "outside_fast_range": lambda: scenes.random_triangles(3000, seed=13, lo=0.0, hi=1e30),
})


@pytest.mark.parametrize("name", list(PACKED_SCENES))
def test_lbvh_build_packed_equals_build_then_pack(gpu, name):
    """rta_lbvh_build_packed writes the traversal layout inside the build's
    own passes; its nodes, indices, triangles and status equal
    rta_lbvh_build's, its tri4 and flags rta_bvh_pack's byte for byte, and
    so does the pair record of every inner node."""
    tris = PACKED_SCENES[name]()
    say(f"packed build scene={name} n={len(tris)}")
    a, b = _build_then_pack(tris), _build_packed(tris)
    assert a["nodes"].tobytes() == b["nodes"].tobytes()
    assert np.array_equal(a["idx"], b["idx"])
    assert a["tris"].tobytes() == b["tris"].tobytes()
    assert a["status"] == b["status"]
    assert np.array_equal(a["tri4"], b["tri4"])
    assert a["flags"] == b["flags"]
    inner = (a["nodes"]["type_left_elems"] & 1) == 1
    assert np.array_equal(a["pairs"][inner], b["pairs"][inner])
    if name == "outside_fast_range":
        assert b["flags"] == 1
    if name == "dup_heavy":
        n = len(tris)
        assert inner.sum() < n - 1               # collapses happened
