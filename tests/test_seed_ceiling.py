"""How much can a hit-distance seed save?  The seeded traversal prunes every
child box whose entry distance exceeds a per-ray bound; with the bound set to
the exact closest-hit distance (a perfect predictor) it must return the same
records, and the node visits it saves bound what any learned seed can gain.

On the heightfield + random-triangle scene with diffuse bounce rays the
saving is small: near-first traversal already meets the closest hit early,
the visits left are boxes the ray crosses before it, and the many rays that
escape the scene have no hit to seed with.  The assertions pin that ceiling
(documented with the full-size numbers elsewhere)."""
from __future__ import annotations

import numpy as np

import oracle
from rta_amd import scenes

F32MAX = np.float32(np.finfo(np.float32).max)


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
    with np.errstate(over="ignore"):
        bound = np.where(hit, h["t"], F32MAX).astype(np.float32)
    hs, _, ss = oracle.trace_kernel(ob["tris"], ob["nodes"], bo, bd, bt, bound=bound, stats=True)
    # the perfect seed changes no record ...
    assert np.array_equal(hs["t"], h["t"]) and np.array_equal(hs["ref"], h["ref"])
    # ... and saves under 10 % of the inner-node visits of the rays it can
    # help (hits), none of the misses'
    v_plain, v_seed = st[hit, 0].mean(), ss[hit, 0].mean()
    assert v_seed <= v_plain
    assert v_seed > 0.9 * v_plain
    assert np.array_equal(ss[~hit, 0], st[~hit, 0])
