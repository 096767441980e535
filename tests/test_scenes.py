"""Synthetic workloads: sizes, determinism and camera rays."""
from __future__ import annotations

import numpy as np

from rta_amd import scenes
from rta_amd.abi import FLT_MAX


def test_icosphere_counts():
    t = scenes.icosphere(5)
    assert len(t) == 20480
    r = np.linalg.norm(t["a"].astype(np.float64), axis=1)
    assert np.allclose(r, 1.0, atol=1e-6)
    assert (t["material_index"] == -1).all()
    verts = np.unique(np.concatenate([t["a"], t["b"], t["c"]]), axis=0)
    assert len(verts) == 10242


def test_heightfield_and_terrain_size():
    hf = scenes.heightfield(64, 32, seed=1)
    assert len(hf) == 2 * 64 * 32
    z = np.concatenate([hf["a"][:, 2], hf["b"][:, 2], hf["c"][:, 2]])
    assert np.abs(z).max() <= 0.2 + 1e-6 and z.std() > 0.01
    assert len(scenes.terrain_scene(grid=16, n_random=10)) == 2 * 16 * 16 + 10


def test_generators_are_deterministic():
    a = scenes.terrain_scene(grid=32, n_random=100)
    b = scenes.terrain_scene(grid=32, n_random=100)
    assert a.tobytes() == b.tobytes()
    assert scenes.random_triangles(50, seed=1).tobytes() != scenes.random_triangles(50, seed=2).tobytes()


def test_random_triangle_edges():
    t = scenes.random_triangles(2000, seed=5)
    c = (t["a"] + t["b"] + t["c"]) / 3
    assert c.min() > -0.02 and c.max() < 1.02
    e = np.linalg.norm(t["a"] - t["b"], axis=1)
    assert e.max() <= 0.02 + 1e-6


def test_camera_rays():
    cam = scenes.sphere_camera(64, 48)
    o, d, t = cam.host_rays()
    assert o.shape == (3 * 64 * 48,) and (t == FLT_MAX).all()
    dd = d.reshape(-1, 3).astype(np.float64)
    assert np.allclose(np.linalg.norm(dd, axis=1), 1.0, atol=1e-6)
    # the centre ray looks at the target
    assert dd[:, 2].min() < -0.8


def test_instances_sizes_and_determinism():
    for n_inst, g in ((1, 2), (5, 8), (16, 32)):
        v, i = scenes.instances(n_inst, g, seed=3)
        assert len(i) == 3 * n_inst * 2 * g * g
        assert i.max() < len(v) // 3
        v2, i2 = scenes.instances(n_inst, g, seed=3)
        assert np.array_equal(v, v2) and np.array_equal(i, i2)
    v, i = scenes.instances(16, 32, seed=3)
    v = v.reshape(-1, 3)
    # instances stay inside their unit cells of a 4 x 4 grid
    assert v[:, 0].min() >= 0 and v[:, 0].max() <= 4 and v[:, 1].max() <= 4
    # every vertex is used and no triangle is degenerate in its ids
    t = i.reshape(-1, 3)
    assert len(np.unique(t)) == len(v)
    assert ((t[:, 0] != t[:, 1]) & (t[:, 1] != t[:, 2]) & (t[:, 0] != t[:, 2])).all()


def test_instances_mix_spheres_and_patches():
    """Each instance is either a geodesic sphere (vertices on a sphere) or a
    heightfield patch (vertices on a regular x/y lattice); both kinds occur."""
    n_inst, g = 16, 16
    v, i = scenes.instances(n_inst, g, seed=5)
    v = v.reshape(-1, 3).astype(np.float64)
    t = i.reshape(n_inst, -1)
    kinds = []
    for k in range(n_inst):
        ids = np.unique(t[k])
        p = v[ids]
        c = p.mean(axis=0)
        r = np.linalg.norm(p - c, axis=1)
        if len(ids) == g * g + 2 and np.ptp(r) < 1e-5 * r.mean():
            kinds.append("sphere")
            assert 0.3 - 1e-6 <= r.mean() <= 0.4 + 1e-6
        else:
            assert len(ids) == (g + 1) ** 2
            xs, ys = np.unique(np.round(p[:, 0] * g)), np.unique(np.round(p[:, 1] * g))
            assert len(xs) == g + 1 and len(ys) == g + 1
            assert np.abs(p[:, 2]).max() <= 0.2 + 1e-6
            kinds.append("patch")
    assert "sphere" in kinds and "patch" in kinds


def test_jitter_statistics_and_shared_vertices():
    import oracle
    v, i = scenes.instances(4, 16, seed=7)
    sigma = 1e-3
    a = oracle.jitter_triangles(v, i, sigma, seed=9, frame=0)
    b = oracle.jitter_triangles(v, i, sigma, seed=9, frame=1)
    t = i.reshape(-1, 3)
    pos = {}
    for tri, rec in zip(t[:2000], a[:2000]):
        for vid, key in zip(tri, ("a", "b", "c")):
            p = tuple(rec[key])
            assert pos.setdefault(int(vid), p) == p     # a shared vertex moves once
    d = np.concatenate([a["a"] - v.reshape(-1, 3)[t[:, 0]], a["b"] - v.reshape(-1, 3)[t[:, 1]]])
    assert np.abs(d.mean(axis=0)).max() < 5e-5
    assert np.allclose(d.std(axis=0), sigma, rtol=0.03)
    assert np.abs(d).max() <= 2 * np.sqrt(3) * sigma * (1 + 1e-5)   # Irwin-Hall(4) support
    assert (a["a"] != b["a"]).mean() > 0.99                       # frames differ
