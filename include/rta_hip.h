/*
 * rta_hip.h -- C-ABI of the bbvh-hip GPU BVH path for MI355X (gfx950).
 *
 * This is the drop-in boundary below the bbvh-hip plugin: plain pointers,
 * sizes and an explicit stream, no C++ or framework types.  Every entry
 * point returns an int status: 0 on success, a negative RTA_E_* code for
 * argument / state errors detected on the host, or a positive HIP runtime
 * error code.  rta_hip_last_error() gives a human readable message for the
 * last failure on the calling thread.  Nothing in this library exits the
 * process (the reference prints and calls exit(), librta/cuda.h:70-78).
 *
 * Device pointers are plain addresses from rta_hip_malloc (or any hipMalloc
 * allocation of the same process).  Launches are asynchronous on the given
 * stream; NULL means the legacy default stream.
 *
 * Data layouts are byte-compatible with the reference:
 *   rta_triangle  104 B  librta/basic_types.h:67-74, librta/cuda-vec.h:30-37
 *   rta_aabb       24 B  librta/cuda-vec.h:40-43
 *   rta_node       32 B  bbvh/bbvh-node.h:5-43 (== bbvh_node_float4 byte
 *                        image, bbvh-cuda/bbvh-cuda-node.h:37-91)
 *   rta_hit        16 B  librta/basic_types.h:81-95
 * Rays are three arrays: origin[3R], direction[3R] (packed float3) and
 * max_t[R] (librta/cuda.h:90-107).
 */
#ifndef RTA_HIP_H
#define RTA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

typedef struct rta_float3 { float x, y, z; } rta_float3;
typedef struct rta_float2 { float x, y; } rta_float2;

typedef struct __attribute__((aligned(8))) rta_triangle {
	rta_float3 a, b, c;
	rta_float3 na, nb, nc;
	rta_float2 ta, tb, tc;
	int32_t material_index;
} rta_triangle;

typedef struct rta_aabb { rta_float3 min, max; } rta_aabb;

/* type_left_elems: bit0 = inner, >>1 = left child (inner) or triangle count
 * (leaf); right_tris: right child (inner) or first triangle (leaf). */
typedef struct rta_node {
	uint32_t type_left_elems;
	uint32_t right_tris;
	rta_aabb box;
} rta_node;

typedef struct rta_hit { float t, beta, gamma; uint32_t ref; } rta_hit;

/* Derived traversal layout: one 64-byte record per inner node holding both
 * children's link words and boxes (SoA within the record), so an inner node
 * visit is one aligned 64-byte fetch.  Built losslessly from rta_node[]. */
typedef struct __attribute__((aligned(64))) rta_node_pair {
	float lo_x[2], lo_y[2], lo_z[2];
	float hi_x[2], hi_y[2], hi_z[2];
	uint32_t link[2];   /* child word: (id << 1) | 1 for an inner child,
	                       first triangle << 1 for a leaf              */
	uint32_t index[2];  /* child node ids                               */
} rta_node_pair;

/* Per-stage build times in milliseconds (hipEvent based), in the stage
 * order of bbvh-cuda/lbvh.h:173-198.  The hierarchy is emitted by the
 * bottom-up pass, so `tree` is (close to) zero and `boxes` covers both. */
typedef struct rta_lbvh_times {
	float init_triangles, voxel_bounds, morton, sort, tree, boxes, permute;
	float total;
} rta_lbvh_times;

typedef void *rta_stream;
typedef void *rta_event;
typedef void *rta_graph;
typedef void *rta_batch;

enum {
	RTA_OK = 0,
	RTA_E_INVALID = -1,     /* bad argument (NULL pointer, n <= 0, ...)  */
	RTA_E_WORKSPACE = -2,   /* workspace smaller than required           */
	RTA_E_NOT_READY = -3,   /* library not initialised / no device       */
	RTA_E_STACK = -4        /* traversal stack overflow reported by GPU  */
};

/* Build status bits written by rta_lbvh_build into *status (device word). */
enum {
	RTA_LBVH_CODE_ZERO = 1u, /* some Morton code is 0: the reference's
	                            bottom-up pass is racy there
	                            (bbvh-cuda/lbvh-kernels.cu:288-291)       */
	RTA_LBVH_FRONTIER_OVERFLOW = 2u
	                         /* a bottom-up frontier list (128 subtree
	                            roots per 1024-leaf chunk) overflowed, so
	                            some upper boxes were not merged.  Cannot
	                            happen: each tree level lengthens the
	                            common prefix of the 64-bit (code,
	                            position) keys, so a tree has at most 64
	                            inner levels and a chunk at most 2 frontier
	                            roots per level.  Reported, not assumed.  */
};

/* ---------------------------------------------------------------- runtime */

const char *rta_hip_last_error(void);
/* Error-path testing (no counterpart in the reference, whose CUDA failures
 * print and exit, librta/cuda.h:70-78): the next call of `entry` returns
 * `status` (non-zero) before doing any work, with the message "<entry>:
 * injected fault"; one shot.  Entry points with this hook: rta_hip_malloc,
 * rta_lbvh_build, rta_lbvh_build_packed, rta_trace_cis_persistent,
 * rta_trace_cis_persistent_ovf, rta_pred_infer, rta_pred_train_step,
 * rta_pred_train_batches. */
int rta_hip_inject_fault(const char *entry, int status);
/* Named host-side ranges (roctx) for a profiler's timeline; the library
 * marks rta_lbvh_build and the persistent traversals with them. */
int rta_hip_range_push(const char *name);
int rta_hip_range_pop(void);
int rta_hip_device_count(int *count);
int rta_hip_set_device(int device);
int rta_hip_device_name(char *buf, int buflen);
int rta_hip_malloc(void **ptr, size_t bytes);
int rta_hip_free(void *ptr);
int rta_hip_memcpy_htod(void *dst, const void *src, size_t bytes);
int rta_hip_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int rta_hip_memcpy_dtod(void *dst, const void *src, size_t bytes, rta_stream s);
/* Fill with a kernel (not hipMemset), safe inside captured graphs. */
int rta_hip_fill_u32(uint32_t *dst, uint32_t value, size_t count, rta_stream s);
/* *word += delta on the device (one thread), e.g. a frame counter. */
int rta_hip_add_u32(uint32_t *word, uint32_t delta, rta_stream s);
int rta_hip_stream_create(rta_stream *s);
int rta_hip_stream_destroy(rta_stream s);
int rta_hip_stream_synchronize(rta_stream s);
int rta_hip_device_synchronize(void);
int rta_hip_event_create(rta_event *e);
int rta_hip_event_destroy(rta_event e);
int rta_hip_event_record(rta_event e, rta_stream s);
/* Work enqueued on s after this call waits for the event's last record
 * (hipStreamWaitEvent); inside a capture it joins s to the capture. */
int rta_hip_stream_wait_event(rta_stream s, rta_event e);
int rta_hip_event_synchronize(rta_event e);
int rta_hip_event_elapsed_ms(float *ms, rta_event start, rta_event stop);
/* Stream capture into a replayable graph (hipStreamBeginCapture, thread-local
 * mode): every entry point of this library that takes a stream enqueues
 * only kernels and copies on it, never a memset or a host synchronisation
 * (the build's optional stage timers excepted), so a whole frame can be
 * captured once and replayed with rta_hip_graph_launch.  s must be a
 * created stream (not NULL).  Buffers are baked into the graph: replays
 * reuse the same addresses and sizes.  No counterpart in the reference
 * (whose launches are synchronous). */
int rta_hip_graph_begin(rta_stream s);
int rta_hip_graph_end(rta_stream s, rta_graph *g);
int rta_hip_graph_launch(rta_graph g, rta_stream s);
int rta_hip_graph_destroy(rta_graph g);
/* Plain coalesced copy kernel of `bytes` (multiple of 16), used to calibrate
 * HBM traffic counters against a kernel of known size. */
int rta_hip_copy_kernel(void *dst, const void *src, size_t bytes, rta_stream s);

/* Self-test of the division-free slab arithmetic on the current device:
 * compares the FMA-corrected quotient used by rta_trace_cis_packed with the
 * GPU's IEEE division on n seeded operand pairs of the admitted ranges.
 * *mismatches gets the count; example (2 floats, may be NULL) one failing
 * (a, b).  No counterpart in the reference (which divides directly). */
int rta_hip_check_division(uint64_t n, uint64_t seed, uint64_t *mismatches,
                           float *example);

/* Enumeration of the one-correction quotient used by the traversal kernels:
 * for every divisor significand B = 1 + mb*2^-23 with mb in
 * [mb_begin, mb_begin + mb_count) and every dividend significand A in [1, 2),
 * compares it with the GPU's IEEE quotient A/B.  mb_begin + mb_count <= 2^23
 * (the 64 calls of 2^17 cover all 2^46 pairs, about a second each).
 * example (2 uint32 bit patterns, may be NULL) gets one failing (A, B).
 * Synchronous.  No counterpart in the reference. */
int rta_hip_check_division_pairs(uint32_t mb_begin, uint32_t mb_count, uint64_t *mismatches,
                                 uint32_t *example);

/* ---------------------------------------------- LBVH stages (Karras 2012) */

/* initTriangles, bbvh-cuda/lbvh-kernels.cu:344-359,482-493:
 * scratch[i] = src[i]; boxes[i] = AABB(tri i); centers[i] = ((a+b)+c)/3.
 * Also writes one partial scene box per 256 triangles into `partials`
 * (ceil(n/256) entries) for rta_lbvh_voxel_bounds. */
int rta_lbvh_init_triangles(rta_triangle *scratch, const rta_triangle *src,
                            rta_aabb *boxes, rta_float3 *centers,
                            rta_aabb *partials, int n, rta_stream s);

/* calculateVoxelBounds, bbvh-cuda/lbvh-kernels.cu:518-535: min/max over the
 * per-block partials (initial box +-1e11f).  The result stays on device. */
int rta_lbvh_voxel_bounds(rta_aabb *bounds, const rta_aabb *partials,
                          int n, rta_stream s);

/* calculateMortonCodes2 + k::morton_codes2, lbvh-kernels.cu:169-183,405-421:
 * d = 1024/(max-min) per axis, q = (uint)((c-min)*d) & 0x3ff, 30-bit
 * interleave (x in bit 0); indices[i] = i.  ORs RTA_LBVH_CODE_ZERO into
 * *status (device word, may be NULL) when a code is 0. */
int rta_lbvh_morton_codes(uint32_t *codes, uint32_t *indices,
                          const rta_float3 *centers, const rta_aabb *bounds,
                          uint32_t *status, int n, rta_stream s);

/* sortMortonCodes, lbvh-kernels.cu:537-555: stable ascending sort of the
 * 30-bit codes carrying indices (hand-written LDS radix sort). */
int rta_lbvh_sort_work_bytes(int n, size_t *bytes);
int rta_lbvh_sort(uint32_t *codes, uint32_t *indices, int n,
                  void *work, size_t work_bytes, rta_stream s);

/* Scratch shared by the two calls below (visit counters, node ranges,
 * first-arrival records, frontier lists): rta_lbvh_boxes_work_bytes(n)
 * bytes.  The reference allocates its counters inside the timed call and
 * never frees them (lbvh-kernels.cu:451-453); here the caller owns them. */
int rta_lbvh_boxes_work_bytes(int n, size_t *bytes);

/* buildTreeKarras + k::build_tree_karras, lbvh-kernels.cu:186-256,424-446:
 * inner nodes 0..n-2, leaves n-1..2n-2; writes links and parents[child]
 * (and the inner nodes' ranges into `work`).  rta_lbvh_build and
 * rta_lbvh_build_packed do not call it: their bottom-up pass emits the
 * same links itself. */
int rta_lbvh_build_tree(rta_node *nodes, uint32_t *parents, void *work,
                        const uint32_t *codes, int n, rta_stream s);

/* buildBoundingBoxesKarras2 + k::build_bounding_boxes_karras2,
 * lbvh-kernels.cu:258-341,448-478: bottom-up box union with a visit
 * counter per inner node; a parent whose two children are leaves with
 * equal representative codes collapses into one multi-triangle leaf.
 * The climb derives the hierarchy from the sorted codes itself (the radix
 * tree over (code, position) is unique) and writes every node whole, with
 * the same links rta_lbvh_build_tree writes, so `parents` is not read;
 * `work`: rta_lbvh_boxes_work_bytes(n) bytes. */
int rta_lbvh_build_boxes(rta_node *nodes, void *work,
                         const rta_aabb *boxes, const uint32_t *codes,
                         const uint32_t *indices, const uint32_t *parents,
                         int n, rta_stream s);

/* invert_triangle_permutation_cuda, lbvh-kernels.cu:557-578:
 * out[i] = scratch[indices[i]]. */
int rta_lbvh_permute(rta_triangle *out, const rta_triangle *scratch,
                     const uint32_t *indices, int n, rta_stream s);

/* lbvh_constructor::buildCUDA, bbvh-cuda/lbvh.h:153-211: all seven stages.
 * tris (device, n) is reordered in place into Morton order (it is the
 * acceleration structure's triangle array afterwards); nodes (2n-1) and
 * indices (n, sorted position -> input index) are outputs.  `status` is an
 * optional device word (RTA_LBVH_* bits).  times may be NULL; when given,
 * the call synchronises the stream to read the stage events. */
int rta_lbvh_workspace_bytes(int n, size_t *bytes);
int rta_lbvh_build(rta_triangle *tris, int n, rta_node *nodes,
                   uint32_t *indices, void *workspace, size_t workspace_bytes,
                   uint32_t *status, rta_lbvh_times *times, rta_stream s);

/* rta_lbvh_build that also writes the traversal layout rta_bvh_pack derives
 * (pairs: 2n-1 records, of which those of inner nodes are written; tri4:
 * 3n float4; flags: one word), with the same bytes, inside the build's own
 * passes: the permutation writes tri4 and runs before the bottom-up pass,
 * which writes each inner node's pair record as it completes the node.
 * times then holds the permutation before the boxes.  No counterpart in the
 * reference (which traces its 2f4 node array directly). */
int rta_lbvh_build_packed(rta_triangle *tris, int n, rta_node *nodes, uint32_t *indices,
                          rta_node_pair *pairs, float *tri4, uint32_t *flags, void *workspace,
                          size_t workspace_bytes, uint32_t *status, rta_lbvh_times *times,
                          rta_stream s);

/* rta_lbvh_build without a workspace argument, for callers that cannot
 * allocate device memory (code under graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same stages, outputs and status bits as rta_lbvh_build; no stage timers.
 * The workspace is a fixed 46 MiB device array inside the library, one per
 * GPU, held until the process ends; the call allocates nothing.  n above
 * RTA_LBVH_NOALLOC_MAX_TRIS returns RTA_E_WORKSPACE before any work.  All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU
 * (other streams, other host threads) race; calls in one stream do not. */
#define RTA_LBVH_NOALLOC_MAX_TRIS 262144
int rta_lbvh_build_noalloc(rta_triangle *tris, int n, rta_node *nodes, uint32_t *indices,
                           uint32_t *status, rta_stream s);

/* --------------------------------------------------------------- tracing */

/* reset_intersections, librta/cuda-kernels.cu:24-44: t = FLT_MAX, beta = 1,
 * gamma = 2, ref untouched, for w*h hits. */
int rta_reset_intersections(rta_hit *hits, int w, int h, rta_stream s);

/* setup_shirley_rays, librta/cuda-kernels.cu:50-112 (pinhole camera of
 * librta/raytrav.h:308-351).  The camera basis and tan(fovy/2) are formed on
 * the host once, the per-pixel arithmetic on the GPU is the host
 * generator's, so the rays are bit-identical to rta_host_shirley_rays. */
int rta_setup_shirley_rays(float *dirs, float *origins, float *max_t,
                           float fovy, float aspect, int w, int h,
                           const float *view_dir, const float *pos,
                           const float *up, float maxt, rta_stream s);

/* trace_cis (2f4 CIS) + launcher, bbvh-cuda/kernels.cu:777-855,1217-1226:
 * closest hit over the 32-byte node array, child-is, near child first
 * (left wins ties), strict t < closest.t and t <= max_t.  Rays 0..w*h-1;
 * hits must hold the reset sentinel (rta_reset_intersections): a miss keeps
 * it.  stats (optional, 2 words per ray) receives the number of inner-node
 * visits and triangle tests of each ray.  err (optional device word) counts
 * stack overflows (RTA_E_STACK on the host side when non-zero). */
int rta_trace_cis(const rta_triangle *tris, int n_tris, const rta_node *nodes,
                  const float *ray_orig, const float *ray_dir,
                  const float *max_t, int w, int h, rta_hit *hits,
                  uint32_t *stats, uint32_t *err, rta_stream s);

/* Derived layout for the fast path: pairs[i] for inner node i (n_nodes
 * records, only inner nodes' written) and the vertex-only triangle array
 * tri4 (3 float4 per triangle: a, a-b, a-c with the same float values
 * intersect_tri_opt forms; the w of a leaf's last triangle is nonzero).
 * flags (device word): bit 0 set when some box coordinate is outside the
 * range where the division-free slab test is exact (then the traversal
 * uses IEEE divisions for every ray). */
int rta_bvh_pack(rta_node_pair *pairs, float *tri4, uint32_t *flags,
                 const rta_node *nodes, int n_nodes,
                 const rta_triangle *tris, int n_tris, rta_stream s);

/* Same semantics and results as rta_trace_cis over the packed layout.
 * flags: the word rta_bvh_pack wrote.  root_link/root_index: the root
 * node's two link words (host copies of nodes[0].type_left_elems /
 * right_tris).  n_rays bounds the launch; when ray_count (device word) is
 * given, only rays below min(n_rays, *ray_count) are traced, so a ray count
 * produced on the GPU (rta_diffuse_bounce) needs no host round trip. */
int rta_trace_cis_packed(const float *tri4, const rta_node_pair *pairs,
                         const uint32_t *flags,
                         uint32_t root_link, uint32_t root_index,
                         const float *ray_orig, const float *ray_dir,
                         const float *max_t, int n_rays,
                         const uint32_t *ray_count, rta_hit *hits,
                         uint32_t *err, rta_stream s);

/* Launch shape of the persistent tracer: block size (64, 128 or 256
 * threads), refill_min (a wave fetches new rays once at least this many of
 * its 64 lanes are idle, 1..64), steps (traversal steps between refill
 * checks), blocks_per_cu (0 = as many as fit) and leaf_min (a wave tests
 * leaf triangles once at least ceil(leaf_min * active / 64) of its active
 * lanes wait at a leaf, or no lane is at an inner node; 1..64, 0 = default
 * 10) and stack (LDS stack entries: 32, the reference's depth, or 16 or 24
 * for a smaller LDS footprint and more resident waves, with an overflow
 * list (the _ovf entry points and the batch API): a ray that would need
 * more is traced again with 32 entries; 0 = 32).  None of these changes a
 * result. */
typedef struct rta_trace_config {
	int block, refill_min, steps, blocks_per_cu, leaf_min, stack;
} rta_trace_config;

/* Persistent form of rta_trace_cis_packed: a machine-filling grid of waves
 * that refill idle lanes from a queue counter (`queue`, one device word of
 * scratch, cleared by this call with a kernel).  Same results, except that
 * a miss only stores t = FLT_MAX and leaves beta, gamma and ref as they
 * were (the reset sentinel).  cfg may be NULL (defaults 64/16/32/0/10). */
int rta_trace_cis_persistent(const float *tri4, const rta_node_pair *pairs,
                             const uint32_t *flags, uint32_t root_link,
                             uint32_t root_index, const float *ray_orig,
                             const float *ray_dir, const float *max_t,
                             int n_rays, const uint32_t *ray_count,
                             rta_hit *hits, uint32_t *err, uint32_t *queue,
                             const rta_trace_config *cfg, rta_stream s);

/* trace_shadow_cis, bbvh-cuda/kernels.cu:857-939 (the `-s` tracer,
 * bbvh-cuda/main.cpp:93): the closest-hit traversal up to the first
 * triangle a leaf accepts (t < closest.t && t <= max_t, closest.t starting
 * at FLT_MAX); that triangle's record is stored and the ray ends.  A miss
 * stores t = FLT_MAX over the reset sentinel.  Same arguments as the
 * closest-hit forms. */
int rta_trace_shadow_cis_packed(const float *tri4, const rta_node_pair *pairs,
                                const uint32_t *flags, uint32_t root_link,
                                uint32_t root_index, const float *ray_orig,
                                const float *ray_dir, const float *max_t,
                                int n_rays, const uint32_t *ray_count,
                                rta_hit *hits, uint32_t *err, rta_stream s);
int rta_trace_shadow_cis_persistent(const float *tri4,
                                    const rta_node_pair *pairs,
                                    const uint32_t *flags, uint32_t root_link,
                                    uint32_t root_index, const float *ray_orig,
                                    const float *ray_dir, const float *max_t,
                                    int n_rays, const uint32_t *ray_count,
                                    rta_hit *hits, uint32_t *err,
                                    uint32_t *queue,
                                    const rta_trace_config *cfg, rta_stream s);

/* The two persistent traversals with a short LDS stack (no counterpart in
 * the reference, whose stack depth is fixed at 32, bbvh-cuda/kernels.cu:
 * 787): with cfg->stack 16 or 24 a ray that would push beyond it is
 * abandoned onto `overflow` (device words: a count, then up to n_rays ray
 * ids; cleared by the call with a kernel) and traced again with 32 entries
 * by a second launch over that list, so every record equals the 32-entry
 * traversal's.  Fewer entries mean more resident waves: coherent rays gain,
 * incoherent ones can lose L2 hits.  With cfg->stack 32 (or cfg NULL) the
 * list is not used; overflow must not be NULL either way. */
int rta_trace_cis_persistent_ovf(const float *tri4, const rta_node_pair *pairs,
                                 const uint32_t *flags, uint32_t root_link,
                                 uint32_t root_index, const float *ray_orig,
                                 const float *ray_dir, const float *max_t,
                                 int n_rays, const uint32_t *ray_count,
                                 rta_hit *hits, uint32_t *err, uint32_t *queue,
                                 const rta_trace_config *cfg,
                                 uint32_t *overflow, rta_stream s);
int rta_trace_shadow_cis_persistent_ovf(const float *tri4,
                                        const rta_node_pair *pairs,
                                        const uint32_t *flags,
                                        uint32_t root_link,
                                        uint32_t root_index,
                                        const float *ray_orig,
                                        const float *ray_dir,
                                        const float *max_t, int n_rays,
                                        const uint32_t *ray_count,
                                        rta_hit *hits, uint32_t *err,
                                        uint32_t *queue,
                                        const rta_trace_config *cfg,
                                        uint32_t *overflow, rta_stream s);

/* ------------------------------------------------ batches of many meshes */

/* A batch: up to 256 independent meshes (1 .. 2^24-1 triangles each), each
 * with its own rays, built and traced together: every stage of the LBVH
 * build is one launch over all meshes and all rays are traced by one
 * persistent launch.  No counterpart in the reference (one mesh per
 * plugin instance); every mesh's results equal a separate build/trace of it.
 *
 * Layouts (mesh m, offsets from rta_batch_offsets):
 *   triangles  concatenated, mesh m at tri_off[m]; the build reorders each
 *              mesh's range in place, as rta_lbvh_build does;
 *   nodes      mesh m's 2n-1 nodes at node_off[m], with child and triangle
 *              indices local to the mesh (exactly rta_lbvh_build's array);
 *   indices    mesh m's sorted->original map at tri_off[m], local;
 *   pairs      mesh m's pair records at node_off[m]; tri4 records at
 *              3*tri_off[m] (float4 units); flags one word per mesh;
 *   rays/hits  mesh m's rays at ray_off[m]; hit refs local to the mesh.
 * rta_batch_create copies the size tables to the device (synchronously);
 * n_rays may be NULL (no rays). */
int rta_batch_create(int n_meshes, const int *n_tris, const int *n_rays, rta_batch *batch);
int rta_batch_destroy(rta_batch batch);
int rta_batch_sizes(rta_batch batch, long long *total_tris, long long *total_nodes,
                    long long *total_rays);
int rta_batch_offsets(rta_batch batch, uint32_t *tri_off, uint32_t *node_off,
                      uint32_t *ray_off);
int rta_batch_workspace_bytes(rta_batch batch, size_t *bytes);
/* All seven build stages for every mesh (status as for rta_lbvh_build). */
int rta_batch_build(rta_batch batch, rta_triangle *tris, rta_node *nodes, uint32_t *indices,
                    void *workspace, size_t workspace_bytes, uint32_t *status, rta_stream s);
/* rta_bvh_pack of every mesh; also records each mesh's root word. */
int rta_batch_pack(rta_batch batch, rta_node_pair *pairs, float *tri4, uint32_t *flags,
                   const rta_node *nodes, const rta_triangle *tris, rta_stream s);
/* rta_batch_build with the traversal layout written by the build's own
 * passes (as rta_lbvh_build_packed does for one mesh): the same nodes,
 * indices, triangles and status, and the same tri4, flags, root words and
 * inner nodes' pair records as rta_batch_pack after rta_batch_build. */
int rta_batch_build_packed(rta_batch batch, rta_triangle *tris, rta_node *nodes,
                           uint32_t *indices, rta_node_pair *pairs, float *tri4,
                           uint32_t *flags, void *workspace, size_t workspace_bytes,
                           uint32_t *status, rta_stream s);
/* Closest-hit (any_hit = 0) or shadow (any_hit = 1) traversal of all rays
 * of all meshes in one persistent launch; results as the single-mesh
 * persistent forms.  Reset the hits first. */
int rta_batch_trace_persistent(rta_batch batch, const float *tri4, const rta_node_pair *pairs,
                               const uint32_t *flags, const float *ray_orig, const float *ray_dir,
                               const float *max_t, rta_hit *hits, uint32_t *err, uint32_t *queue,
                               int any_hit, const rta_trace_config *cfg, rta_stream s);
/* Rays the last batch traversal with stack 16 or 24 traced again because
 * they needed more entries (synchronous; 0 after a 32-entry traversal). */
int rta_batch_overflow_count(rta_batch batch, uint32_t *count, rta_stream s);
/* Autotuner of the persistent launch shape on this batch: coordinate
 * descent over leaf_min, refill_min (the queue chunk a wave takes), steps,
 * block and stack, `rounds` rounds, each candidate timed as the best of `reps`
 * launches (synchronous).  *best: the starting shape (block 0 = defaults);
 * on return the fastest found and *best_ms its time.  The hits of the last
 * launch are valid results. */
int rta_batch_autotune(rta_batch batch, const float *tri4, const rta_node_pair *pairs,
                       const uint32_t *flags, const float *ray_orig, const float *ray_dir,
                       const float *max_t, rta_hit *hits, uint32_t *err, uint32_t *queue,
                       int any_hit, int rounds, int reps, rta_trace_config *best,
                       float *best_ms, rta_stream s);

/* -------------------------------------------------------------- bouncing */

/* One cosine-weighted diffuse bounce per hit (the incoherent ray set).
 * For every ray i with hits[i].t != FLT_MAX, in ray order, writes a bounce
 * ray: origin = o + t*d + 1e-4*n, n the unit geometric normal of triangle
 * hits[i].ref facing the incoming ray, direction = normalize(n + s) with s
 * uniform on the unit sphere (Marsaglia 1972, counter-based RNG keyed on
 * (seed, i)), max_t = FLT_MAX.  src_index[k] = i.  *count (device word) =
 * number of bounce rays.  work: rta_bounce_work_bytes(n_rays). */
int rta_bounce_work_bytes(int n_rays, size_t *bytes);
int rta_diffuse_bounce(float *out_orig, float *out_dir, float *out_maxt,
                       uint32_t *src_index, uint32_t *count,
                       const rta_hit *hits, const rta_triangle *tris,
                       const float *ray_orig, const float *ray_dir,
                       int n_rays, uint64_t seed, void *work,
                       size_t work_bytes, rta_stream s);

/* Per-frame geometry of a dynamic indexed mesh (rta_scene_instances): writes
 * n_tris triangle records from verts (3 floats per vertex) and idx (3 vertex
 * ids per triangle), every vertex moved by sigma times a unit-variance
 * Irwin-Hall(4) variate per axis, keyed on (seed, frame, vertex id), so
 * shared vertices move together.  Normals and texcoords 0, material -1.
 * frame_word (device word, may be NULL) replaces `frame` when given, so a
 * captured frame can advance it with rta_hip_add_u32.  No counterpart in
 * the reference. */
int rta_jitter_triangles(rta_triangle *out, const float *verts, const uint32_t *idx,
                         int n_tris, float sigma, uint64_t seed, uint32_t frame,
                         const uint32_t *frame_word, rta_stream s);

/* One shadow ray per hit towards a point light (light: host float[3]):
 * origin = o + t*d + 1e-4*n as for the bounce, direction = the unit vector
 * from that origin to the light, max_t = the distance to it (a light at the
 * origin itself gives direction n and max_t 0).  Compaction, src_index,
 * *count and work as for rta_diffuse_bounce.  No counterpart in the
 * reference (whose bouncers trace no further rays, librta/cuda.h:232-234). */
int rta_shadow_rays(float *out_orig, float *out_dir, float *out_maxt,
                    uint32_t *src_index, uint32_t *count, const rta_hit *hits,
                    const rta_triangle *tris, const float *ray_orig,
                    const float *ray_dir, int n_rays, const float *light,
                    void *work, size_t work_bytes, rta_stream s);

/* ------------------------------------------- learned hit predictor (MLP)
 * No counterpart in the reference.  A small network proposes, per ray, the
 * treelets most likely to hold its closest hit (treelet c = entry c of a
 * fixed cut of the tree into at most 1024 subtrees, rta_treelet_cut) and a
 * log hit distance; rta_trace_seed_treelets turns the proposals into a
 * max_t seed and the exact traversal decides every result.  Network: 102
 * encoded inputs
 * (origin mapped to [-1,1] by the scene box, unit direction, sin/cos of
 * 2^f*pi*v for f = 0..7), three ReLU layers of 64, 1024 treelet logits + 1
 * log distance; bf16 weights and activations, FP32 accumulation (MFMA),
 * FP32 master weights and Adam.  Parameters are one packed FP32 array of
 * rta_pred_param_count floats: W1[64][112], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[1040][64], b4[1040] (weights out-major; padding
 * columns/rows stay zero). */
typedef void *rta_pred;
int rta_pred_param_count(int *count);
/* State for batches of up to max_batch rays, He-uniform weights from seed. */
int rta_pred_create(int max_batch, uint64_t seed, rta_pred *pred);
int rta_pred_destroy(rta_pred pred);
int rta_pred_get_params(rta_pred pred, float *host);
int rta_pred_set_params(rta_pred pred, const float *host);
int rta_pred_get_grads(rta_pred pred, float *host);
int rta_pred_zero_grads(rta_pred pred, rta_stream s);
/* Forward pass of n rays (device arrays; scene_min/max: 3 host floats).
 * out (n x 1040 f32), top (n x 4 treelet ids, best first) and logdist (n)
 * are optional device outputs.  Keeps the activations for a backward pass. */
int rta_pred_forward(rta_pred pred, const float *ray_orig, const float *ray_dir, int n,
                     const float *scene_min, const float *scene_max, float *out,
                     int32_t *top, float *logdist, rta_stream s);
/* Gradients of the last forward pass (added to the gradient buffer):
 * softmax cross-entropy of the 1024 logits against label[i] (a treelet, or
 * -1 for a ray without hit, which contributes nothing) plus the squared
 * error of the log distance against target[i], averaged over the n rays;
 * the loss is added to *loss (device float). */
int rta_pred_backward(rta_pred pred, const int32_t *label, const float *target, int n,
                      float *loss, rta_stream s);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the accumulated
 * gradients; step counts from 1. */
int rta_pred_adam(rta_pred pred, float lr, int step, rta_stream s);
/* The same with the step count in device memory: *step is incremented and
 * the step then taken with it (for captured, replayed launch sequences). */
int rta_pred_adam_device(rta_pred pred, float lr, uint32_t *step, rta_stream s);
/* One fused training step on rays [0, n) (n in 1..max_batch): forward pass,
 * the loss of rta_pred_backward, gradients and an Adam step in three
 * launches, without atomics, so the result does not depend on scheduling.
 * The 1040 outputs of a ray are never stored.  The gradient buffer
 * (rta_pred_get_grads) receives this step's gradient, replacing its
 * contents; *step (device word) is incremented and the Adam step taken with
 * it, as rta_pred_adam_device does; the loss is added to *loss (device
 * float).  Bias sums of the 1024 logits' softmax and of every gradient run
 * in another order than rta_pred_forward/backward, so results agree with
 * that path to float rounding, not bit for bit. */
int rta_pred_train_step(rta_pred pred, const float *ray_orig, const float *ray_dir,
                        const int32_t *label, const float *target, int n,
                        const float *scene_min, const float *scene_max, float lr,
                        uint32_t *step, float *loss, rta_stream s);
/* rta_pred_train_step over n_batches batches of the n rays: batch b covers
 * rays [starts[b], min(starts[b] + batch, n)) (starts: host array, each in
 * [0, n); batch in 1..max_batch).  The launches of all batches are queued
 * by one call. */
int rta_pred_train_batches(rta_pred pred, const float *ray_orig, const float *ray_dir,
                           const int32_t *label, const float *target, int n,
                           const int32_t *starts, int n_batches, int batch,
                           const float *scene_min, const float *scene_max, float lr,
                           uint32_t *step, float *loss, rta_stream s);
/* n_out rays with their hit records drawn from the first *count rays of a
 * set (device arrays): entry i takes ray hash(seed, frame, i) mod *count,
 * frame = *frame_word (or 0 when frame_word is NULL); an empty set gives
 * miss records.  The training sample of an online fine-tune. */
int rta_pred_sample(const float *ray_orig, const float *ray_dir, const rta_hit *hits,
                    const uint32_t *count, int n_out, const uint32_t *frame_word, uint64_t seed,
                    float *orig_out, float *dir_out, rta_hit *hits_out, rta_stream s);

/* Fused inference for n rays (no batch limit): encoding, the four layers
 * and the top-k selection in one kernel, the logits kept in registers.
 * top[4*i..] (optional): the top_k (4 or 1) treelets with the largest keys,
 * largest first, where a treelet's key is its logit with the low 10
 * mantissa bits replaced by 1023 - treelet (with top_k 1, entries 1..3 are
 * -1); logdist[i] (optional): the log-distance output; logits (optional,
 * n x 1040 floats, for testing): all outputs.  The weights are those of the
 * last rta_pred_set_params / Adam step on the same stream.  Same network
 * and encoding as rta_pred_forward; the sums run in another order (16x16x32
 * tiles, bias first), so logits may differ from it in the last bits. */
int rta_pred_infer(rta_pred pred, const float *ray_orig, const float *ray_dir, int n,
                   const float *scene_min, const float *scene_max, int32_t *top, int top_k,
                   float *logdist, float *logits, rta_stream s);

/* Karras ranges (sorted positions [from, to], 2 words per inner node, n-1
 * nodes) of the last rta_lbvh_build made with this workspace. */
int rta_lbvh_copy_ranges(uint32_t *ranges, const void *workspace, int n, rta_stream s);

/* One treelet of a cut: the subtree rooted at `word` (a pair-record word:
 * inner node id<<1|1, or a leaf's first triangle<<1) covering the sorted
 * triangle positions [first, next treelet's first); box = its root's box
 * (the whole tree's root: +-FLT_MAX). */
typedef struct rta_treelet {
	float lo[3], hi[3];
	uint32_t word, first;
} rta_treelet;

/* A fixed cut of the packed tree into at most max_treelets (1..1024)
 * subtrees, found in rounds: starting from the root, each round takes the
 * inner treelets holding more than half as many triangles as the largest
 * inner one and splits them, largest first (ties to the lower first
 * position), as many as fit, into their two children.  table
 * (max_treelets entries) receives the treelets in ascending `first` (so
 * they partition [0, n_tris)), *count (device word) their number.  ranges:
 * rta_lbvh_copy_ranges of the same build.  One workgroup. */
int rta_treelet_cut(const rta_node_pair *pairs, uint32_t root_link, uint32_t root_index,
                    const uint32_t *ranges, int n_tris, int max_treelets, rta_treelet *table,
                    uint32_t *count, rta_stream s);

/* Training labels from closest-hit records: label[i] = the treelet whose
 * range holds hits[i].ref and target[i] = log(hits[i].t), or -1 and 0 for a
 * miss (t = FLT_MAX). */
int rta_treelet_labels(const rta_hit *hits, int n, const rta_treelet *table,
                       const uint32_t *count, int32_t *label, float *target, rta_stream s);

/* Hit-distance seeds from predicted treelets: for each ray, the closest hit
 * among the triangles of its first `candidates` (1..4) treelets of
 * top[4*i..] (ids >= *count are skipped), traversed from their roots in
 * order of root-box entry distance; t_seed[i] = that t, or FLT_MAX.  Every
 * such t is the t of a real hit of the ray (the same triangle test,
 * t <= max_t), so the closest hit lies at or before it. */
int rta_trace_seed_treelets(const float *tri4, const rta_node_pair *pairs, const uint32_t *flags,
                            const rta_treelet *table, const uint32_t *count,
                            const float *ray_orig, const float *ray_dir, const float *max_t,
                            int n_rays, const int32_t *top, int candidates, float *t_seed,
                            uint32_t *err, rta_stream s);
/* The seeding and the exact traversal in one persistent launch: ray i first
 * traverses treelet top[4*i] (when it is < *count and the ray meets its root
 * box within max_t[i]) up to the first leaf that holds a hit, then the whole
 * tree with max_t' = min(max_t[i], s + s * 2^-10), s the smallest t of that
 * phase, from a closest hit of FLT_MAX, as rta_trace_cis_persistent.  Any
 * hit of the ray bounds its closest one; the 2^-10 margin (as in
 * rta_seed_max_t) covers boxes whose computed entry distance exceeds the
 * computed t of a triangle inside them, so the records are the plain
 * traversal's. */
int rta_trace_cis_persistent_seeded(const float *tri4, const rta_node_pair *pairs,
                                    const uint32_t *flags, uint32_t root_link,
                                    uint32_t root_index, const float *ray_orig,
                                    const float *ray_dir, const float *max_t, int n_rays,
                                    const uint32_t *ray_count, const rta_treelet *table,
                                    const uint32_t *count, const int32_t *top, rta_hit *hits,
                                    uint32_t *err, uint32_t *queue, const rta_trace_config *cfg,
                                    rta_stream s);
/* out[i] = min(max_t[i], t_seed[i] + t_seed[i] * 2^-10): the max_t a seeded
 * traversal runs with.  The seed (the t of a real hit) is widened by 2^-10
 * relative because a box holding the closest triangle can have a computed
 * entry distance slightly beyond that triangle's computed t; pruning at the
 * seed itself could lose the closest hit. */
int rta_seed_max_t(float *out, const float *max_t, const float *t_seed, int n, rta_stream s);

#ifdef __cplusplus
}
#endif

#endif /* RTA_HIP_H */
