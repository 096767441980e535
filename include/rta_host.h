/*
 * rta_host.h -- host-side helpers of the bbvh-hip path (plain C ABI):
 * seeded synthetic scene generators and the host pinhole ray generator.
 *
 * All generators are deterministic functions of their arguments (counter
 * based hashing, no global RNG state), so every run and every process sees
 * byte-identical scenes and rays.
 */
#ifndef RTA_HOST_H
#define RTA_HOST_H

#include <stdint.h>
#include "rta_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Icosphere: regular icosahedron subdivided `subdiv` times (shared
 * vertices), projected to `radius` around `center`.  20*4^subdiv triangles;
 * normals are the normalised positions, texcoords 0, material -1.
 * Returns the triangle count; `out` may be NULL to query it.  `jitter` > 0
 * displaces every shared vertex radially by jitter*U(-1,1) (seeded). */
int rta_scene_icosphere(int subdiv, float radius, const float *center, float jitter,
                        uint64_t seed, rta_triangle *out);

/* Heightfield over [x0,x1]x[y0,y1]: nx*ny quads of two triangles each (fixed
 * diagonal), z = z0 + amplitude * fBm(x, y) where fBm is `octaves` octaves
 * of seeded value noise (base frequency `base_freq` lattice cells per unit,
 * lacunarity, gain), normalised by the octave weight sum (|fBm| <= 1).
 * Returns 2*nx*ny; `out` may be NULL. */
int rta_scene_heightfield(int nx, int ny, float x0, float y0, float x1, float y1, float z0,
                          uint64_t seed, int octaves, float amplitude, float lacunarity,
                          float gain, float base_freq, rta_triangle *out);

/* `count` small triangles: centres uniform in [lo,hi]^3, each vertex at
 * centre + len_k * dir_k with len_k uniform in [edge_min, edge_max] and dir_k
 * uniform on the sphere.  Returns count; `out` may be NULL. */
int rta_scene_random_triangles(int count, uint64_t seed, float lo, float hi, float edge_min,
                               float edge_max, rta_triangle *out);

/* The large dynamic scene as an indexed mesh: n_inst instances on a square
 * grid of unit cells (instance k in cell (k % side, k / side), side =
 * ceil(sqrt(n_inst))), each, by a seeded coin, either a value-noise
 * heightfield patch of g x g quads (6 octaves, amplitude 0.2) or a geodesic
 * sphere (octahedron subdivided until it has 2*g*g triangles, radius
 * 0.3..0.4, resting 0.1 above the ground).  Every instance has 2*g*g
 * triangles (g a power of two >= 2; g = 256 and 64 instances give
 * 8,388,608).  verts: 3 floats per vertex; idx: 3 vertex ids per triangle,
 * instance by instance.  Returns the triangle count and sets *n_verts; with
 * verts or idx NULL nothing else is written. */
int rta_scene_instances(int n_inst, int g, uint64_t seed, int *n_verts, float *verts,
                        uint32_t *idx);

/* Host version of the pinhole generator (librta/raytrav.h:300-351): the
 * same camera basis and per-pixel arithmetic as rta_setup_shirley_rays, so
 * host and device rays are bit-identical.  Ray id = y*w + x. */
int rta_host_shirley_rays(float *dirs, float *origins, float *max_t, float fovy, float aspect,
                          int w, int h, const float *view_dir, const float *pos, const float *up,
                          float maxt);

/* Spatial-median binary BVH on the host (bbvh_constructor_using_median with
 * spatial_median, bbvh/bbvh.h:110-279), the tree the bbvh-cuda plugin builds
 * on the CPU and uploads for `-b median` (bbvh-cuda/main.cpp:158-163,
 * 200-235).  nodes: room for rta_host_median_node_bound(n) = 2n-1 nodes,
 * filled in pre-order (32-byte rta_node, leaves of at most 4 triangles);
 * tris_out (optional, n): the triangles in leaf order, the array the tree's
 * leaves index; order (optional, n): the input index of each of them.
 * Returns the node count, or -1 for n <= 0 or a NULL tris/nodes.  Node boxes,
 * split axes, split positions and the order among equal centroid
 * components are the reference's (finite input). */
int rta_host_median_node_bound(int n);
int rta_host_median_build(const rta_triangle *tris, int n, rta_node *nodes, rta_triangle *tris_out,
                          uint32_t *order);

/* ---- the reference's measurement protocol (rta/main.cpp:200-462) ----
 * Functions below return 0 (or a count) on success and a negative value on
 * failure, with the message in rta_host_last_error(). */
const char *rta_host_last_error(void);

/* Bounds of all triangle vertices as compute_aabb forms them
 * (librta/aabb.h:10-49): start from the first triangle's vertex a, update
 * on strict < / >.  n >= 1. */
int rta_host_scene_bounds(const rta_triangle *tris, int n, float *bb_min, float *bb_max);

/* Camera of one spherical-series sample (rta/main.cpp:338-346, 355-373):
 * pos = centre + normalize(vertex) * |max-min| * distance_factor, with
 * centre = min + (max-min)/2; dir = normalize(-vertex) (towards the
 * coordinate origin); up from (0,0,1), or (0,1,0) when |dot(dir, z)| > 0.8,
 * then up = cross(dir, cross(dir, up)).  The series renders with fovy 45. */
int rta_host_series_camera(const float *vertex, const float *bb_min, const float *bb_max,
                           float distance_factor, float *pos, float *dir, float *up);

/* Camera positions of the axial series (rta/main.cpp:296-309): sample i is
 * the point (anchor.x, anchor.x, anchor.x) -- the reference builds it from
 * the x component three times -- rotated about `axis` (normalised here) by
 * float((2 * 3.1416 / samples) * i) radians.  out: 3 floats per sample. */
int rta_host_axial_positions(const float *axis, const float *anchor, int samples, float *out);

/* Distinct vertex positions of an ASCII .ply file (x, y, z properties of
 * element "vertex"), first occurrence first, as rta/main.cpp:247-270 keeps
 * them.  Returns the count; at most max_out are written (out may be NULL). */
int rta_host_ply_vertices(const char *path, float *out, int max_out);

/* Triangles of a Wavefront .obj file: v/vt/vn records, faces with v,
 * v/t, v//n or v/t/n corners (negative indices count back), polygons split
 * into fans around their first corner.  Normals and texcoords are taken
 * where given (else 0); material_index is -1.  Returns the triangle count;
 * at most max_out are written (out may be NULL). */
int rta_host_load_obj(const char *path, rta_triangle *out, int max_out);

#ifdef __cplusplus
}
#endif

#endif /* RTA_HOST_H */
