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


def test_heightfield_and_config2_size():
    hf = scenes.heightfield(64, 32, seed=1)
    assert len(hf) == 2 * 64 * 32
    z = np.concatenate([hf["a"][:, 2], hf["b"][:, 2], hf["c"][:, 2]])
    assert np.abs(z).max() <= 0.2 + 1e-6 and z.std() > 0.01
    assert len(scenes.config2_scene(grid=16, n_random=10)) == 2 * 16 * 16 + 10


def test_generators_are_deterministic():
    a = scenes.config2_scene(grid=32, n_random=100)
    b = scenes.config2_scene(grid=32, n_random=100)
    assert a.tobytes() == b.tobytes()
    assert scenes.random_triangles(50, seed=1).tobytes() != scenes.random_triangles(50, seed=2).tobytes()


def test_random_triangle_edges():
    t = scenes.random_triangles(2000, seed=5)
    c = (t["a"] + t["b"] + t["c"]) / 3
    assert c.min() > -0.02 and c.max() < 1.02
    e = np.linalg.norm(t["a"] - t["b"], axis=1)
    assert e.max() <= 0.02 + 1e-6


def test_camera_rays():
    cam = scenes.config1_camera(64, 48)
    o, d, t = cam.host_rays()
    assert o.shape == (3 * 64 * 48,) and (t == FLT_MAX).all()
    dd = d.reshape(-1, 3).astype(np.float64)
    assert np.allclose(np.linalg.norm(dd, axis=1), 1.0, atol=1e-6)
    # the centre ray looks at the target
    assert dd[:, 2].min() < -0.8
