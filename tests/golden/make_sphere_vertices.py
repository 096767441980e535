"""Writes tests/golden/sphere0_vertices.json: the distinct vertex positions
of the reference's measurement sphere sphs/sphere0.ply (the camera sphere of
bbvh-cuda-bench.sh), in file order, as the decimal strings the file holds.

Reads the ASCII .ply directly (header, then one row per vertex with x y z
first); runs only where the reference tree is present:
    python tests/golden/make_sphere_vertices.py [/root/reference/sphs/sphere0.ply]
"""
import json
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/sphs/sphere0.ply"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere0_vertices.json")


def main():
    with open(SRC) as f:
        lines = f.read().splitlines()
    n_vertex, props, i = None, [], 1
    while lines[i] != "end_header":
        parts = lines[i].split()
        if parts[0] == "element":
            cur = parts[1]
            if cur == "vertex":
                n_vertex = int(parts[2])
        elif parts[0] == "property" and cur == "vertex":
            props.append(parts[-1])
        i += 1
    ix, iy, iz = props.index("x"), props.index("y"), props.index("z")
    rows = [lines[i + 1 + k].split() for k in range(n_vertex)]
    seen, distinct = set(), []
    for r in rows:
        key = (float(r[ix]), float(r[iy]), float(r[iz]))
        if key not in seen:
            seen.add(key)
            distinct.append([r[ix], r[iy], r[iz]])
    with open(OUT, "w") as f:
        json.dump({"source": "sphs/sphere0.ply", "rows": n_vertex, "distinct": distinct}, f,
                  indent=0)
    print(f"{len(distinct)} distinct of {n_vertex} vertices -> {OUT}")


if __name__ == "__main__":
    main()
