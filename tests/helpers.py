"""Independent checks used to pin the oracle (and, through it, the GPU path).

Nothing here restates the oracle's procedures: the radix tree is built by a
recursive split on the highest differing bit of the augmented keys, boxes
and collapsed leaves come from their range characterisation, and triangle
answers from exact rational arithmetic or float64 linear algebra.
"""
from __future__ import annotations

import sys
from fractions import Fraction

import numpy as np

sys.setrecursionlimit(10000)


def morton_bitloop(x: int, y: int, z: int) -> int:
    """Bit i of x -> bit 3i, of y -> 3i+1, of z -> 3i+2 (low 10 bits)."""
    code = 0
    for i in range(10):
        code |= ((x >> i) & 1) << (3 * i)
        code |= ((y >> i) & 1) << (3 * i + 1)
        code |= ((z >> i) & 1) << (3 * i + 2)
    return code


def radix_tree(codes_sorted: np.ndarray):
    """Binary radix tree over the unique keys (code << 32 | position).

    Returns {inner index: (lo, hi, left, right)} with Karras numbering: the
    root is 0 and covers [0, n-1]; a node covering [lo, hi] split after s has
    left child s (leaf n-1+s if s == lo) and right child s+1 (leaf n-1+hi if
    s+1 == hi)."""
    n = len(codes_sorted)
    keys = (codes_sorted.astype(np.uint64) << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    keys = [int(k) for k in keys]
    out = {}
    stack = [(0, 0, n - 1)]
    while stack:
        idx, lo, hi = stack.pop()
        b = (keys[lo] ^ keys[hi]).bit_length() - 1
        # keys are sorted: bit b is 0 on a prefix of the range, 1 after it
        a, z = lo, hi
        while a < z:  # last position with bit b == 0
            m = (a + z + 1) // 2
            if (keys[m] >> b) & 1:
                z = m - 1
            else:
                a = m
        s = a
        left = s if s > lo else n - 1 + s
        right = s + 1 if s + 1 < hi else n - 1 + hi
        out[idx] = (lo, hi, left, right)
        if s > lo:
            stack.append((s, lo, s))
        if s + 1 < hi:
            stack.append((s + 1, s + 1, hi))
    return out


def expected_lbvh_nodes(codes_sorted, indices, tri_boxes_min, tri_boxes_max):
    """Links and boxes every node must have, from the range characterisation:
    a node's box is the AABB of the triangles in its sorted range; an inner
    node becomes a leaf (tris = lo, elems = hi-lo+1) iff all codes in its
    range are equal."""
    n = len(codes_sorted)
    tree = radix_tree(codes_sorted) if n > 1 else {}
    nn = 2 * n - 1
    tle = np.zeros(nn, np.uint32)
    rt = np.zeros(nn, np.uint32)
    bmin = np.zeros((nn, 3), np.float32)
    bmax = np.zeros((nn, 3), np.float32)
    smin = tri_boxes_min[indices]
    smax = tri_boxes_max[indices]
    for k in range(n):
        tle[n - 1 + k] = 2
        rt[n - 1 + k] = k
        bmin[n - 1 + k] = smin[k]
        bmax[n - 1 + k] = smax[k]
    for idx, (lo, hi, left, right) in tree.items():
        bmin[idx] = smin[lo:hi + 1].min(axis=0)
        bmax[idx] = smax[lo:hi + 1].max(axis=0)
        if codes_sorted[lo] == codes_sorted[hi]:
            tle[idx] = (hi - lo + 1) << 1
            rt[idx] = lo
        else:
            tle[idx] = (left << 1) | 1
            rt[idx] = right
    parents = np.full(nn, 0xFFFFFFFF, np.uint32)
    for idx, (lo, hi, left, right) in tree.items():
        parents[left] = idx
        parents[right] = idx
    return tle, rt, bmin, bmax, parents


def round_f32(fr: Fraction) -> np.float32:
    """Correctly rounded (nearest-even) float32 of an exact rational."""
    f = np.float32(float(fr))
    best = None
    for cand in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - fr)
        key = (err, int(np.frombuffer(np.float32(cand).tobytes(), np.uint32)[0]) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return np.float32(best[1])


def exact_cramer(a, b, c, o, d):
    """Exact (t, beta, gamma) of o + t d = a + beta (b-a) + gamma (c-a)."""
    F = lambda v: [Fraction(float(x)) for x in v]
    a, b, c, o, d = F(a), F(b), F(c), F(o), F(d)
    # columns: (a-b), (a-c), d ; rhs a-o  (Cramer with the same orientation)
    m = [[a[i] - b[i], a[i] - c[i], d[i]] for i in range(3)]
    r = [a[i] - o[i] for i in range(3)]

    def det(mm):
        return (mm[0][0] * (mm[1][1] * mm[2][2] - mm[1][2] * mm[2][1])
                - mm[0][1] * (mm[1][0] * mm[2][2] - mm[1][2] * mm[2][0])
                + mm[0][2] * (mm[1][0] * mm[2][1] - mm[1][1] * mm[2][0]))

    D = det(m)
    if D == 0:
        return None
    sol = []
    for col in range(3):
        mc = [row[:] for row in m]
        for i in range(3):
            mc[i][col] = r[i]
        sol.append(det(mc) / D)
    beta, gamma, t = sol
    return t, beta, gamma


def tri_boxes(tris):
    v = np.stack([tris["a"], tris["b"], tris["c"]], axis=1)
    return v.min(axis=1), v.max(axis=1)


# ------------------------------------------------ planar scenes for seeding

def planar_scene(n_random: int = 30_000, seed: int = 5, z: float = 0.3):
    """Random small triangles (edges 1e-4..0.02) all lying in the plane
    z = `z`, plus a flat 64x64 grid in the same plane: every leaf box has
    zero thickness in z, so its entry distance comes from one slab quotient
    while a triangle's t comes from Cramer's rule, and the two round
    differently."""
    from rta_amd import scenes
    r = scenes.random_triangles(n_random, seed=seed, edge_min=1e-4, edge_max=0.02)
    for k in ("a", "b", "c"):
        r[k][:, 2] = np.float32(z)
    g = scenes.heightfield(64, 64, z0=z, amplitude=0.0, seed=1)
    return np.concatenate([r, g])


def grazing_rays(n: int, seed: int = 1, z: float = 0.3):
    """Rays from just above the plane z = `z`, heading down at slopes of
    1e-4..1 (many of them nearly parallel to the plane)."""
    rng = np.random.default_rng(seed)
    o = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n),
                         np.full(n, z) + rng.uniform(1e-4, 1e-2, n)]).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[:, 2] = -np.abs(d[:, 2]) * rng.choice([1e-4, 1e-3, 1e-2, 0.1, 1.0], n)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o.reshape(-1), d.reshape(-1), np.full(n, np.finfo(np.float32).max, np.float32)


# ------------------------------------------ the reference's text dump formats

def fmt_g(x) -> str:
    """A float as `std::ostream << float` writes it by default (6 significant
    digits, %g)."""
    return "%g" % float(x)


def bvh_dump_lines(nodes: np.ndarray) -> list[str]:
    """The .bvh text of a node array (bbvh/bbvh.h:36-52): header, then per
    node its index, I/L, the box, and split axis (0), left, right for an
    inner node or the count and triangle indices for a leaf."""
    out = ["bvh", str(len(nodes))]
    for i, nd in enumerate(nodes):
        w, r = int(nd["type_left_elems"]), int(nd["right_tris"])
        box = " ".join(fmt_g(v) for v in (*nd["min"], *nd["max"]))
        if w & 1:
            out.append(f"{i} I {box} 0 {w >> 1} {r}")
        else:
            e = w >> 1
            out.append(" ".join([f"{i} L {box} {e}"] + [str(r + j) for j in range(e)]))
    return out


def tri_dump_lines(tris: np.ndarray) -> list[str]:
    """The .tri text of a triangle array (bbvh/bbvh.h:53-65): header, then
    per triangle its index, vertices a, b, c and normals na, nb, nc."""
    out = ["tri", str(len(tris))]
    for i, t in enumerate(tris):
        vals = (*t["a"], *t["b"], *t["c"], *t["na"], *t["nb"], *t["nc"])
        out.append(" ".join([str(i)] + [fmt_g(v) for v in vals]))
    return out
