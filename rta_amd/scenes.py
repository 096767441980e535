"""Seeded synthetic scenes, cameras and the workload configurations.

Scene geometry comes from the C++ generators in librta_host.so (see
include/rta_host.h); everything is a deterministic function of the seeds, so
the GPU path and the CPU oracle always see byte-identical inputs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .abi import FLT_MAX, TRI_DTYPE, host, ptr


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def icosphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0), jitter: float = 0.0,
              seed: int = 0) -> np.ndarray:
    lib = host()
    n = lib.rta_scene_icosphere(subdiv, radius, _f3(center), jitter, seed, None)
    out = np.zeros(n, dtype=TRI_DTYPE)
    lib.rta_scene_icosphere(subdiv, radius, _f3(center), jitter, seed, ptr(out))
    return out


def heightfield(nx: int, ny: int, x0=0.0, y0=0.0, x1=1.0, y1=1.0, z0=0.0, seed: int = 1,
                octaves: int = 6, amplitude: float = 0.2, lacunarity: float = 2.0,
                gain: float = 0.5, base_freq: float = 4.0) -> np.ndarray:
    lib = host()
    args = (nx, ny, x0, y0, x1, y1, z0, seed, octaves, amplitude, lacunarity, gain, base_freq)
    n = lib.rta_scene_heightfield(*args, None)
    if n < 0:
        raise ValueError("bad heightfield arguments")
    out = np.zeros(n, dtype=TRI_DTYPE)
    lib.rta_scene_heightfield(*args, ptr(out))
    return out


def random_triangles(count: int, seed: int = 2, lo=0.0, hi=1.0, edge_min=0.001,
                     edge_max=0.01) -> np.ndarray:
    out = np.zeros(count, dtype=TRI_DTYPE)
    host().rta_scene_random_triangles(count, seed, lo, hi, edge_min, edge_max, ptr(out))
    return out


@dataclass
class Camera:
    pos: tuple
    target: tuple
    up: tuple
    fovy: float
    w: int
    h: int

    @property
    def view_dir(self):
        return tuple(float(t) - float(p) for t, p in zip(self.target, self.pos))

    @property
    def aspect(self) -> float:
        return float(np.float32(self.w) / np.float32(self.h))

    def host_rays(self, maxt: float = float(FLT_MAX)):
        """Pixel-centre pinhole rays on the host: (origins[3R], dirs[3R], max_t[R])."""
        n = self.w * self.h
        o = np.empty(3 * n, np.float32)
        d = np.empty(3 * n, np.float32)
        t = np.empty(n, np.float32)
        rc = host().rta_host_shirley_rays(ptr(d), ptr(o), ptr(t), self.fovy, self.aspect, self.w,
                                          self.h, _f3(self.view_dir), _f3(self.pos), _f3(self.up),
                                          maxt)
        if rc != 0:
            raise ValueError("bad camera")
        return o, d, t


# ------------------------------------------------------- configurations

def sphere_scene() -> np.ndarray:
    """Icosphere of radius 1 subdivided 5 times: 20,480 triangles."""
    return icosphere(5, 1.0)


def sphere_camera(w: int = 512, h: int = 512) -> Camera:
    """Pinhole at distance 3 looking at the sphere, 60 degree vertical FOV."""
    return Camera(pos=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovy=60.0,
                  w=w, h=h)


def terrain_scene(grid: int = 1024, n_random: int = 200_000, seed: int = 7) -> np.ndarray:
    """1024x1024 heightfield over [0,1]^2 (6-octave value noise, amplitude 0.2,
    lacunarity 2, gain 0.5) plus 200,000 small random triangles in the unit
    cube: 2,297,152 triangles at full size."""
    hf = heightfield(grid, grid, seed=seed)
    rnd = random_triangles(n_random, seed=seed + 1000)
    return np.concatenate([hf, rnd])


def terrain_camera(w: int = 1920, h: int = 1080) -> Camera:
    """Above the heightfield looking down at ~50 degrees; ~96% of the
    pixel rays hit geometry."""
    return Camera(pos=(0.5, 0.05, 0.6), target=(0.5, 0.55, 0.0), up=(0.0, 0.0, 1.0),
                  fovy=35.0, w=w, h=h)


BOUNCE_SEED = 0x5EED_B0B0
# The hit-predictor workload on the terrain scene: the network trains on the
# diffuse bounce rays of four seeds (about 8.26 M rays at full size) and is
# evaluated on two held-out seeds (about 4.13 M rays; offset 0 is the
# benchmark's own bounce set).  Offsets from BOUNCE_SEED.
PREDICTOR_TRAIN_SEEDS = (100, 101, 102, 103)
PREDICTOR_EVAL_SEEDS = (0, 1)


def instances(n_inst: int = 64, g: int = 256, seed: int = 11):
    """Indexed mesh of the dynamic scene (rta_scene_instances): n_inst
    heightfield patches / geodesic spheres of 2*g*g triangles each on a grid
    of unit cells.  Returns (verts float32[3V], idx uint32[3T])."""
    nv = C.c_int(0)
    nt = host().rta_scene_instances(n_inst, g, seed, C.byref(nv), None, None)
    if nt < 0:
        raise ValueError("rta_scene_instances: bad arguments")
    verts = np.zeros(3 * nv.value, np.float32)
    idx = np.zeros(3 * nt, np.uint32)
    host().rta_scene_instances(n_inst, g, seed, C.byref(nv), ptr(verts), ptr(idx))
    return verts, idx


def dynamic_camera(w: int = 3840, h: int = 2160, n_inst: int = 64) -> Camera:
    """Looking across the instance grid from one side, slightly above it."""
    side = float(int(np.ceil(np.sqrt(n_inst))))
    return Camera(pos=(0.5 * side, -0.35 * side, 0.6 * side), target=(0.5 * side, 0.55 * side, 0.0),
                  up=(0.0, 0.0, 1.0), fovy=50.0, w=w, h=h)


def dynamic_light(n_inst: int = 64):
    """Point light above one corner of the instance grid."""
    side = float(int(np.ceil(np.sqrt(n_inst))))
    return (0.8 * side, 0.2 * side, 0.5 * side)


DYNAMIC_SIGMA = 1e-3
DYNAMIC_SEED = 0xD1A_2024


def batch_meshes(n_meshes: int = 256, seed: int = 3, tris=(1_000, 200_000),
                 rays=(4_096, 262_144)):
    """The many-meshes workload: n_meshes independent meshes whose triangle
    counts are log-uniform in `tris`, each a jittered icosphere (the largest
    subdivision 20*4^s not above the drawn count, radius 1, at the origin)
    or a value-noise heightfield patch over [0,1]^2 (g x g quads with 2g^2
    close to the count), chosen by a seeded coin; each with its own rays,
    log-uniform in `rays`: half from a random camera 3 units outside the
    mesh towards random points of its box, half from uniform origins in the
    box in uniform directions.  Returns (meshes, rays) lists; rays are
    (origins[3r], dirs[3r], max_t[r]) float32."""
    rng = np.random.default_rng(seed)
    meshes, ray_sets = [], []
    for k in range(n_meshes):
        target = float(np.exp(rng.uniform(np.log(tris[0]), np.log(tris[1]))))
        if rng.uniform() < 0.5 and target >= 20:
            s = int(np.floor(np.log(target / 20.0) / np.log(4.0)))
            s = max(0, min(s, 7))
            m = icosphere(s, 1.0, jitter=0.02, seed=seed * 1000 + k)
        else:
            g = max(1, int(round(np.sqrt(target / 2.0))))
            m = heightfield(g, g, seed=seed * 1000 + k)
        meshes.append(m)
        n = int(np.exp(rng.uniform(np.log(rays[0]), np.log(rays[1]))))
        ray_sets.append(mesh_rays(m, n, rng))
    return meshes, ray_sets


def mesh_rays(m: np.ndarray, n: int, rng):
    """n rays for mesh m: half from a random camera 3 units outside the
    mesh's box centre towards random points of the box, half from uniform
    origins in the box in uniform directions; max_t = FLT_MAX."""
    v = np.concatenate([m["a"], m["b"], m["c"]]).astype(np.float64)
    lo, hi = v.min(axis=0), v.max(axis=0)
    n1 = n // 2
    ctr = 0.5 * (lo + hi)
    u = rng.normal(size=3)
    cam = ctr + 3.0 * u / np.linalg.norm(u)
    d1 = rng.uniform(lo, hi, size=(n1, 3)) - cam
    o1 = np.broadcast_to(cam, (n1, 3))
    o2 = rng.uniform(lo, hi, size=(n - n1, 3))
    d2 = rng.normal(size=(n - n1, 3))
    o = np.concatenate([o1, o2]).astype(np.float32)
    d = np.concatenate([d1, d2])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o.reshape(-1), d.reshape(-1), np.full(n, FLT_MAX, np.float32)
