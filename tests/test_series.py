"""The reference's measurement protocol in rta-hip (rta/main.cpp:200-462,
rta/cmdline.cpp:24-157): sphere .ply sample positions, the spherical and
axial camera series, scene bounds, .obj models, the PLY-with-rps output and
the .bvh/.tri/.ray dumps.

The reference's vector helpers live in libmcm, which is not in the reference
tree, so the camera arithmetic is pinned by the properties the protocol
fixes (distances, directions, the up-vector switch and orientation, the
rotation about the axis) rather than by bit-for-bit vectors; the GPU test
pins the whole chain (sphere file -> camera -> rays -> hits) against the
oracle on the host rays of the same camera."""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from rta_amd import abi
from rta_amd.abi import TRI_DTYPE, host, ptr
from helpers import bvh_dump_lines, fmt_g, tri_dump_lines

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "rta_amd", "bin", "rta-hip")
GOLDEN = os.path.join(ROOT, "tests", "golden", "sphere0_vertices.json")
F32MAX = np.float32(np.finfo(np.float32).max)


def run_driver(*args, timeout=300):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)


def sphere0():
    with open(GOLDEN) as f:
        g = json.load(f)
    return g["distinct"]


def write_ply(path, rows, faces=(), extra_first=False):
    """ASCII ply with x y z nx ny nz per vertex; `rows` are string triples."""
    lines = ["ply", "format ascii 1.0", "comment test sphere"]
    face_block = [f"element face {len(faces)}", "property list uchar uint vertex_indices"]
    vert_block = [f"element vertex {len(rows)}"] + [f"property float {p}" for p in
                                                     ("x", "y", "z", "nx", "ny", "nz")]
    lines += (face_block + vert_block) if extra_first else (vert_block + face_block)
    lines.append("end_header")
    fl = [f"{len(f)} " + " ".join(str(i) for i in f) for f in faces]
    vl = [" ".join(r) + " 0 0 1" for r in rows]
    lines += (fl + vl) if extra_first else (vl + fl)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def ply_vertices(path):
    n = host().rta_host_ply_vertices(str(path).encode(), None, 0)
    if n < 0:
        return n, None
    out = np.zeros(3 * max(n, 1), np.float32)
    assert host().rta_host_ply_vertices(str(path).encode(), ptr(out), n) == n
    return n, out[:3 * n].reshape(-1, 3)


def f3(v):
    return np.ascontiguousarray(v, dtype=np.float32)


def camera(vertex, lo, hi, factor):
    pos, d, up = (np.zeros(3, np.float32) for _ in range(3))
    assert host().rta_host_series_camera(ptr(f3(vertex)), ptr(f3(lo)), ptr(f3(hi)), factor,
                                         ptr(pos), ptr(d), ptr(up)) == 0
    return pos.astype(np.float64), d.astype(np.float64), up.astype(np.float64)


# ---------------------------------------------------------------- CPU

def test_sphere0_fixture_has_the_icosphere_shape():
    """sphs/sphere0.ply is a once-subdivided icosahedron: 42 distinct unit
    vectors (12 of the icosahedron, 30 edge midpoints)."""
    v = np.array(sphere0(), dtype=np.float64)
    assert v.shape == (42, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)


def test_ply_reader_keeps_first_occurrences_in_order(tmp_path):
    rows = sphere0()[:8]
    dup = rows + rows[2:5] + rows[:1]          # repeated positions, as per-face exports hold
    for extra_first in (False, True):          # a list element before the vertices too
        p = tmp_path / f"s{int(extra_first)}.ply"
        write_ply(p, dup, faces=[(0, 1, 2), (2, 3, 4, 5)], extra_first=extra_first)
        n, v = ply_vertices(p)
        assert n == 8
        exp = np.array([[np.float32(x) for x in r] for r in rows], np.float32)
        assert np.array_equal(v, exp)


def test_ply_reader_errors(tmp_path):
    assert ply_vertices(tmp_path / "missing.ply")[0] == -1
    bad = tmp_path / "bin.ply"
    bad.write_text("ply\nformat binary_little_endian 1.0\nelement vertex 1\n"
                   "property float x\nproperty float y\nproperty float z\nend_header\n")
    assert ply_vertices(bad)[0] == -2
    assert b"ascii" in host().rta_host_last_error()
    short = tmp_path / "short.ply"
    short.write_text("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\n"
                     "property float y\nproperty float z\nend_header\n0 0 1\n0 1 0\n")
    assert ply_vertices(short)[0] == -2


def test_scene_bounds_are_the_vertex_extent():
    tris = np.zeros(50, TRI_DTYPE)
    rng = np.random.default_rng(3)
    for f in ("a", "b", "c"):
        tris[f] = rng.normal(size=(50, 3)).astype(np.float32)
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    assert host().rta_host_scene_bounds(ptr(tris), 50, ptr(lo), ptr(hi)) == 0
    allv = np.concatenate([tris["a"], tris["b"], tris["c"]])
    assert np.array_equal(lo, allv.min(axis=0)) and np.array_equal(hi, allv.max(axis=0))
    assert host().rta_host_scene_bounds(ptr(tris), 0, ptr(lo), ptr(hi)) < 0


def test_series_camera_geometry():
    """rta/main.cpp:355-373 for every sphere0 vertex: the camera sits at
    factor * diameter from the scene centre along the vertex direction,
    looks at the coordinate origin, and its up vector is
    dir x (dir x up0) with up0 = z, or y when |dir . z| > 0.8."""
    lo, hi = np.array([-1.0, 2.0, 0.5]), np.array([3.0, 4.0, 1.5])
    centre, diam = lo + (hi - lo) / 2, np.linalg.norm(hi - lo)
    for r in sphere0():
        v = np.array([float(x) for x in r])
        pos, d, up = camera(v, lo, hi, 1.5)
        off = pos - centre
        assert abs(np.linalg.norm(off) - 1.5 * diam) < 1e-5 * diam
        assert np.allclose(off / np.linalg.norm(off), v / np.linalg.norm(v), atol=1e-5)
        # towards the origin, not towards the scene centre
        assert np.allclose(d, -v / np.linalg.norm(v), atol=1e-6)
        up0 = np.array([0.0, 1.0, 0.0]) if abs(d[2]) > 0.8 else np.array([0.0, 0.0, 1.0])
        exp = np.cross(d, np.cross(d, up0))
        assert np.allclose(up, exp, atol=1e-6)
        # the double cross points away from up0
        assert np.dot(up, up0) < 0


def test_series_camera_up_switch_threshold():
    lo, hi = np.zeros(3), np.ones(3)
    for dz, up0 in ((0.81, (0, 1, 0)), (0.79, (0, 0, 1)), (-0.81, (0, 1, 0))):
        v = np.array([np.sqrt(1 - dz * dz), 0.0, -dz])     # dir = -v has z component dz
        _, d, up = camera(v, lo, hi, 1.0)
        assert abs(d[2] - dz) < 1e-6
        u0 = np.array(up0, np.float64)
        assert abs(np.dot(up, np.cross(d, u0))) < 1e-6     # up lies in span(dir, up0)
        assert np.dot(up, u0) < 0


def test_axial_positions():
    """rta/main.cpp:296-309: sample i is (a.x, a.x, a.x) rotated about the
    axis by 2*3.1416*i/n; about z that is the textbook planar rotation."""
    n = 7
    out = np.zeros(3 * n, np.float32)
    anchor = f3([2.0, 5.0, 7.0])
    assert host().rta_host_axial_positions(ptr(f3([0, 0, 3])), ptr(anchor), n, ptr(out)) == 0
    p = out.reshape(-1, 3).astype(np.float64)
    assert np.array_equal(out[:3], np.float32([2, 2, 2]))     # x taken three times
    for i in range(n):
        a = np.float32((2 * 3.1416 / n) * i)
        c, s = np.cos(a), np.sin(a)
        assert np.allclose(p[i], [2 * c - 2 * s, 2 * s + 2 * c, 2.0], atol=1e-5)
    # a tilted axis: norm and the component along the axis are kept
    k = np.array([1.0, 2.0, 2.0]) / 3
    assert host().rta_host_axial_positions(ptr(f3([1, 2, 2])), ptr(anchor), n, ptr(out)) == 0
    q = out.reshape(-1, 3).astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), np.sqrt(12), atol=1e-5)
    assert np.allclose(q @ k, np.dot([2, 2, 2], k), atol=1e-5)
    assert host().rta_host_axial_positions(ptr(f3([0, 0, 1])), ptr(anchor), 0, ptr(out)) < 0


def test_obj_loader(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# test\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0.5 0.25\nvn 0 0 1\n"
                 "f 1 2 3 4\n"                 # quad -> fan (1,2,3) (1,3,4)
                 "f -4/1/1 -3/1/1 -2/1/1\n"   # negative indices, v/t/n
                 "f 1//1 3//1 4//1\n")
    n = host().rta_host_load_obj(str(p).encode(), None, 0)
    assert n == 4
    t = np.zeros(n, TRI_DTYPE)
    assert host().rta_host_load_obj(str(p).encode(), ptr(t), n) == n
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    assert np.array_equal(t["a"], V[[0, 0, 0, 0]])
    assert np.array_equal(t["b"], V[[1, 2, 1, 2]])
    assert np.array_equal(t["c"], V[[2, 3, 2, 3]])
    assert np.array_equal(t["na"][:2], np.zeros((2, 3), np.float32))
    assert np.array_equal(t["na"][2:], np.float32([[0, 0, 1]] * 2))
    assert np.array_equal(t["ta"][2], np.float32([0.5, 0.25]))
    assert np.all(t["material_index"] == -1)
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    assert host().rta_host_load_obj(str(bad).encode(), ptr(t), n) == -2
    assert host().rta_host_load_obj(str(tmp_path / "none.obj").encode(), None, 0) == -1


def test_driver_series_argument_checks(tmp_path):
    """Errors the reference's parser raises (rta/cmdline.cpp:146-155) come
    before any device work, so they are checked here without a GPU."""
    r = run_driver("-p", "0,0,3", "--axis", "0,0,1", "-s", "4", timeout=60)
    assert r.returncode == 1 and "exclusive" in r.stderr
    r = run_driver("--axis", "0,0,1", timeout=60)
    assert r.returncode == 1 and "--samples" in r.stderr
    r = run_driver("--sphere-file", str(tmp_path / "none.ply"), timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
    # the abbreviation bbvh-cuda-bench.sh uses
    r = run_driver("--sphere", str(tmp_path / "none.ply"), "--dist", "1.5", timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
    # -f picks one series out of several
    r = run_driver("-p", "0,0,3", "--sphere-file", str(tmp_path / "none.ply"), "-f", "s",
                   timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_driver_sphere_series(gpu, tmp_path):
    """A spherical series over six sphere0 vertices (written with repeats)
    on the icosphere scene: one dot per distinct vertex, the PLY with rps,
    the .bvh/.tri/.ray dumps, and the last camera's hits equal to the
    oracle's on the host rays of the same camera."""
    rows = sphere0()[:6]
    sph = tmp_path / "sphere.ply"
    write_ply(sph, rows + rows[:2], faces=[(0, 1, 2)])
    outp, dump, prefix = tmp_path / "out.ply", tmp_path / "hits.bin", str(tmp_path / "mamo")
    w, h = 160, 120
    r = run_driver("-r", f"{w}x{h}", "--scene", "icosphere:4", "--sphere-file", str(sph),
                   "--distance-factor", "1.5", "-o", str(outp), "-M", prefix,
                   "--dump-hits", str(dump), "bbvh-hip", "--", "-b", "lbvh")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "got 6 distinct vertices" in r.stdout
    assert "......\naverage rps per frame:" in r.stdout
    # PLY with rps: every row of a sampled position carries its timing
    text = outp.read_text().splitlines()
    assert "property float rps" in text and any(l.startswith("comment meta ray-tracer:")
                                                  for l in text)
    start = text.index("end_header") + 1
    rps = [float(l.split()[-1]) for l in text[start:start + 8]]
    assert all(v > 0 for v in rps) and rps[6] == rps[0] and rps[7] == rps[1]
    # dumps
    bvh = open(prefix + ".bvh").read().splitlines()
    tri = open(prefix + ".tri").read().splitlines()
    n = 20 * 4 ** 4
    assert bvh[:2] == ["bvh", str(2 * n - 1)] and tri[:2] == ["tri", str(n)]
    assert all(os.path.exists(f"{prefix}{k:03d}.ray") for k in range(6))
    # the last camera, recomputed on the host
    tris = np.zeros(n, TRI_DTYPE)
    host().rta_scene_icosphere(4, 1.0, None, 0.0, 0, ptr(tris))
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    host().rta_host_scene_bounds(ptr(tris), n, ptr(lo), ptr(hi))
    v = f3([float(x) for x in rows[5]])
    pos, d, up = (np.zeros(3, np.float32) for _ in range(3))
    host().rta_host_series_camera(ptr(v), ptr(lo), ptr(hi), 1.5, ptr(pos), ptr(d), ptr(up))
    R = w * h
    o, dd, t = np.empty(3 * R, np.float32), np.empty(3 * R, np.float32), np.empty(R, np.float32)
    aspect = float(np.float32(w) / np.float32(h))
    host().rta_host_shirley_rays(ptr(dd), ptr(o), ptr(t), 45.0, aspect, w, h, ptr(d), ptr(pos),
                                 ptr(up), float(F32MAX))
    # the .ray body (librta/raytrav.h:353-366): id, 2, origin, direction,
    # 0, max_t, each float as the stream writes it (%g)
    ray = open(f"{prefix}005.ray").read().splitlines()
    assert ray[:2] == ["ray", f"{w} {h}"]
    o3, d3 = o.reshape(-1, 3), dd.reshape(-1, 3)
    exp = [" ".join([str(i), "2"] + [fmt_g(v) for v in (*o3[i], *d3[i])] + ["0", fmt_g(t[i])])
           for i in range(R)]
    assert ray[2:] == exp
    got = np.fromfile(dump, dtype=abi.HIT_DTYPE)
    ob = oracle.lbvh_build(tris)
    # the .bvh/.tri bodies (bbvh/bbvh.h:36-65): the oracle's LBVH nodes and
    # Morton-ordered triangles in the reference's text format
    assert bvh == bvh_dump_lines(ob["nodes"])
    assert tri == tri_dump_lines(ob["tris"])
    ref, _, _ = oracle.trace_cis(ob["tris"], ob["nodes"], o, dd, t, w, h)
    hit = ref["t"] != F32MAX
    assert hit.mean() > 0.05
    assert np.array_equal(got["t"] != F32MAX, hit)
    for f in ("t", "beta", "gamma", "ref"):
        assert np.array_equal(got[f][hit], ref[f][hit])
