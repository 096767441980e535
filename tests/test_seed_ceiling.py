"""How much can a hit-distance seed save, and what bound keeps it exact?

The seeded traversal prunes every child box whose entry distance exceeds a
per-ray bound derived from a seed s, the t of a real hit of the ray.  The
closest hit lies at t <= s, but the computed entry distance of a box that
holds the closest triangle can exceed that triangle's computed t: the slab
test divides once per axis, Cramer's rule rounds differently.  On planar
triangles met at grazing angles the excess reaches about 2^-16.3 relative,
so a bound at s itself loses closest hits there, while the kernels' bound
s + s * 2^-10 (predictor.seed_bound) loses none.

With the bound from the exact closest-hit distance (a perfect predictor)
the records are unchanged, and the node visits it saves bound what any
learned seed can gain.  On the heightfield + random-triangle scene with
diffuse bounce rays the saving is small: near-first traversal already meets
the closest hit early, the visits left are boxes the ray crosses before it,
and the many rays that escape the scene have no hit to seed with."""
from __future__ import annotations

import numpy as np

import oracle
from rta_amd import scenes
from rta_amd.predictor import seed_bound
from helpers import grazing_rays, planar_scene

F32MAX = np.float32(np.finfo(np.float32).max)


def _differs(a, b):
    return (a["ref"] != b["ref"]) | (a["t"].view(np.uint32) != b["t"].view(np.uint32))


def test_seed_bound_is_a_float32_widening():
    s = np.array([1e-30, 1.0, 0.3, 7.5e3, F32MAX], np.float32)
    b = seed_bound(s)
    assert b.dtype == np.float32
    assert np.all(b[:-1] > s[:-1]) and np.isinf(b[-1])
    # 2^-10 relative, rounded once to float32 (an error of at most 2^-24
    # of the sum, i.e. 2^-14 of the 2^-10 excess)
    np.testing.assert_allclose(b[:-1].astype(np.float64) / s[:-1] - 1.0, 2.0 ** -10,
                               rtol=2.0 ** -13)


def test_perfect_seed_saves_little():
    tris = scenes.terrain_scene(grid=256, n_random=20_000)
    ob = oracle.lbvh_build(tris)
    cam = scenes.terrain_camera(320, 180)
    o, d, t = cam.host_rays()
    h0, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], o, d, t)
    bo, bd, bt, _ = oracle.diffuse_bounce(h0, ob["tris"], o, d, scenes.BOUNCE_SEED)
    h, _, st = oracle.trace_kernel(ob["tris"], ob["nodes"], bo, bd, bt, stats=True)
    hit = h["t"] != F32MAX
    assert 0.2 < hit.mean() < 0.9
    bound = seed_bound(np.where(hit, h["t"], F32MAX))
    hs, _, ss = oracle.trace_kernel(ob["tris"], ob["nodes"], bo, bd, bt, bound=bound, stats=True)
    # the perfect seed changes no record ...
    assert not _differs(hs, h).any()
    # ... and saves under 10 % of the inner-node visits of the rays it can
    # help (hits), none of the misses'
    v_plain, v_seed = st[hit, 0].mean(), ss[hit, 0].mean()
    assert v_seed <= v_plain
    assert v_seed > 0.9 * v_plain
    assert np.array_equal(ss[~hit, 0], st[~hit, 0])


def test_planar_scene_needs_the_margin():
    """Planar triangles and grazing rays: a bound at the exact closest t
    loses closest hits (so the scene exercises the rounding gap), one and
    four ulps above it still lose some, and the kernels' 2^-10 bound loses
    none."""
    tris = planar_scene()
    ob = oracle.lbvh_build(tris)
    o, d, t = grazing_rays(200_000)
    h, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], o, d, t)
    hit = h["t"] != F32MAX
    assert hit.sum() > 50_000
    exact = np.where(hit, h["t"], F32MAX).astype(np.float32)
    lost = {}
    for name, bound in (("exact", exact),
                        ("4 ulps", np.where(hit, (exact.view(np.uint32) + 4).view(np.float32),
                                            F32MAX).astype(np.float32)),
                        ("2^-10", seed_bound(exact))):
        hs, _ = oracle.trace_kernel(ob["tris"], ob["nodes"], o, d, t, bound=bound)
        lost[name] = int((_differs(hs, h) & hit).sum())
    assert lost["exact"] > 1000, lost
    assert lost["4 ulps"] > 0, lost
    assert lost["2^-10"] == 0, lost
