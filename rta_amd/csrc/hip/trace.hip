// Closest-hit traversal and the GPU glue around it: intersection reset,
// pinhole ray generation, child-pair packing and the diffuse bouncer.
//
// Traversal semantics are those of the reference's 2f4 CIS kernel
// (bbvh-cuda/kernels.cu:777-855), which for max_t = FLT_MAX equal the CPU
// bbvh child-is tracer (bbvh/bbvh.h:407-455): at an inner node both child
// boxes are tested, a child is kept iff it is hit with dist < closest.t and
// dist <= max_t, the nearer child (left on ties) is entered and the other
// pushed; leaves test their triangles in index order and accept
// t < closest.t && t <= max_t.  Box and triangle tests are the exact ones
// of librta/intersect.h (see common.hpp).
#include "internal.hpp"

#include <algorithm>

namespace rta_hip {

constexpr int kTraceBlock = 128;
#ifndef RTA_TRACE_STACK
#define RTA_TRACE_STACK 32
#endif
constexpr int kStack = RTA_TRACE_STACK;   // the reference's stack depth is 32 (kernels.cu:787)
// Upper bound on loop iterations of one ray: a malformed tree ends in an
// error count instead of a hang.  A valid tree needs at most
// 2*nodes iterations; 2^26 covers every tree this library can hold.
constexpr uint32_t kMaxSteps = 1u << 26;

// ------------------------------------------------------------- reset

__global__ void reset_kernel(rta_hit *hits, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	hits[i].t = FLT_MAX;
	hits[i].beta = 1.0f;
	hits[i].gamma = 2.0f;
}

// ---------------------------------------------------------- ray gen

struct CamBasis {
	float ux, uy, uz, vx, vy, vz, wx, wy, wz;
	float px, py, pz;
	float width, height;
};

__global__ void shirley_kernel(float *dirs, float *orgs, float *maxts, int w, int h, CamBasis cb,
                               float maxt) {
	const int x = blockIdx.x * blockDim.x + threadIdx.x;
	const int y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x >= w || y >= h) return;
	const size_t id = (size_t)y * w + x;
	orgs[3 * id + 0] = cb.px;
	orgs[3 * id + 1] = cb.py;
	orgs[3 * id + 2] = cb.pz;
	maxts[id] = maxt;
	float u_s = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
	float v_s = ((float)y + 0.5f) / (float)h * 2.0f - 1.0f;
	u_s = cb.width * u_s;
	v_s = cb.height * v_s;
	float dx = cb.ux * u_s, dy = cb.uy * u_s, dz = cb.uz * u_s;
	dx = dx + cb.vx * v_s;
	dy = dy + cb.vy * v_s;
	dz = dz + cb.vz * v_s;
	dx = dx + cb.wx;
	dy = dy + cb.wy;
	dz = dz + cb.wz;
	const float len = sqrtf(dx * dx + dy * dy + dz * dz);
	dirs[3 * id + 0] = dx / len;
	dirs[3 * id + 1] = dy / len;
	dirs[3 * id + 2] = dz / len;
}

// -------------------------------------------- plain CIS over rta_node[]

// The per-lane stack lives in LDS, entry-major so that a wave's pushes and
// pops at different depths spread over the banks.
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) void trace_cis_kernel(
    const rta_triangle *__restrict__ tris, uint32_t n_tris, const rta_node *__restrict__ nodes,
    uint32_t n_nodes, const float *__restrict__ ro, const float *__restrict__ rd,
    const float *__restrict__ rmax, int n_rays, rta_hit *__restrict__ hits,
    uint32_t *__restrict__ stats, uint32_t *__restrict__ err) {
	__shared__ uint32_t stack[kStack][kTraceBlock];
	const int lt = threadIdx.x;
	const int tid = blockIdx.x * kTraceBlock + lt;
	if (tid >= n_rays) return;

	const float ox = ro[3 * tid + 0], oy = ro[3 * tid + 1], oz = ro[3 * tid + 2];
	const float dx = rd[3 * tid + 0], dy = rd[3 * tid + 1], dz = rd[3 * tid + 2];
	const float tmax = rmax[tid];
	rta_hit closest = hits[tid];
	closest.t = FLT_MAX;
	uint32_t n_inner = 0, n_tested = 0;

	const float4 *nv = reinterpret_cast<const float4 *>(nodes);
	float4 curr = nv[0];
	int sp = -1;
	for (uint32_t guard = 0;; ++guard) {
		const uint32_t cl = __float_as_uint(curr.x);
		const uint32_t cr = __float_as_uint(curr.y);
		if (guard >= kMaxSteps) { atomicAdd(err, 1u); break; }
		if (link_inner(cl)) {
			if (STATS) ++n_inner;
			const uint32_t L = cl >> 1, R = cr;
			if (L >= n_nodes || R >= n_nodes) { atomicAdd(err, 1u); break; }
			const float4 l0 = nv[2 * L], l1 = nv[2 * L + 1];
			const float4 r0 = nv[2 * R], r1 = nv[2 * R + 1];
			float dl, dr;
			bool hl = box_test(l0.z, l0.w, l1.x, l1.y, l1.z, l1.w, ox, oy, oz, dx, dy, dz, dl);
			bool hr = box_test(r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, ox, oy, oz, dx, dy, dz, dr);
			hl = hl && dl < closest.t && dl <= tmax;
			hr = hr && dr < closest.t && dr <= tmax;
			if (hl && hr) {
				if (sp + 1 >= kStack) { atomicAdd(err, 1u); break; }
				if (dl <= dr) { stack[++sp][lt] = R; curr = l0; }
				else          { stack[++sp][lt] = L; curr = r0; }
				continue;
			}
			if (hl) { curr = l0; continue; }
			if (hr) { curr = r0; continue; }
		} else {
			const uint32_t elems = cl >> 1, off = cr;
			if (off >= n_tris || elems > n_tris - off) { atomicAdd(err, 1u); break; }
			for (uint32_t i = 0; i < elems; ++i) {
				if (STATS) ++n_tested;
				const float *tf = reinterpret_cast<const float *>(tris + off + i);
				const float ax = tf[0], ay = tf[1], az = tf[2];
				float t, be, ga;
				if (tri_test(ax, ay, az, ax - tf[3], ay - tf[4], az - tf[5], ax - tf[6],
				             ay - tf[7], az - tf[8], ox, oy, oz, dx, dy, dz, t, be, ga)) {
					if (t < closest.t && t <= tmax) {
						closest.t = t;
						closest.beta = be;
						closest.gamma = ga;
						closest.ref = off + i;
					}
				}
			}
		}
		if (sp < 0) break;
		curr = nv[2 * stack[sp--][lt]];
	}
	hits[tid] = closest;
	if (STATS) {
		stats[2 * tid + 0] = n_inner;
		stats[2 * tid + 1] = n_tested;
	}
}

// ------------------------------------------------ child-pair packing

// word for a child as seen from its parent's pair record:
//   inner child c -> (c << 1) | 1        (index of its own pair record)
//   leaf child    -> (first_tri << 1)    (the leaf's last triangle carries
//                                         a flag in the tri4 array)
__device__ inline uint32_t pair_word(const rta_node &c, uint32_t id) {
	return link_inner(c.type_left_elems) ? ((id << 1) | 1u) : (c.right_tris << 1);
}

__device__ inline void pack_tri(float4 *__restrict__ tri4, const rta_triangle *__restrict__ tris,
                                int i) {
	const float *f = reinterpret_cast<const float *>(tris + i);
	const float ax = f[0], ay = f[1], az = f[2];
	tri4[3 * i + 0] = make_float4(ax, ay, az, 0.0f);
	tri4[3 * i + 1] = make_float4(ax - f[3], ay - f[4], az - f[5], 0.0f);
	tri4[3 * i + 2] = make_float4(ax - f[6], ay - f[7], az - f[8], 0.0f);
}

__global__ void pack_tris_kernel(float4 *__restrict__ tri4, const rta_triangle *__restrict__ tris,
                                 uint32_t *__restrict__ flags, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0) *flags = 0u;
	if (i >= n) return;
	pack_tri(tri4, tris, i);
}

__device__ inline void mark_leaf_end(float4 *tri4, const rta_node &leaf) {
	const uint32_t last = leaf.right_tris + (leaf.type_left_elems >> 1) - 1u;
	tri4[3 * last].w = __uint_as_float(1u);
}

__device__ inline bool box_fast_ok(const rta_aabb &b) {
	return coord_fast_ok(b.min.x) && coord_fast_ok(b.min.y) && coord_fast_ok(b.min.z) &&
	       coord_fast_ok(b.max.x) && coord_fast_ok(b.max.y) && coord_fast_ok(b.max.z);
}

// One pair record per inner node; flags bit 0 is set when some tested box
// has a coordinate outside the fast-safe range (then every ray uses the
// IEEE division path).
__device__ inline void pack_pair(rta_node_pair *__restrict__ pairs, float4 *__restrict__ tri4,
                                 uint32_t *__restrict__ flags, const rta_node *__restrict__ nodes,
                                 int x);

__global__ void pack_pairs_kernel(rta_node_pair *__restrict__ pairs, float4 *__restrict__ tri4,
                                  uint32_t *__restrict__ flags, const rta_node *__restrict__ nodes,
                                  int n_nodes) {
	const int x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= n_nodes) return;
	pack_pair(pairs, tri4, flags, nodes, x);
}

__device__ inline void pack_pair(rta_node_pair *__restrict__ pairs, float4 *__restrict__ tri4,
                                 uint32_t *__restrict__ flags, const rta_node *__restrict__ nodes,
                                 int x) {
	const rta_node nx = nodes[x];
	if (!link_inner(nx.type_left_elems)) {
		if (x == 0) mark_leaf_end(tri4, nx);  // a leaf root has no parent record
		return;
	}
	const uint32_t L = nx.type_left_elems >> 1, R = nx.right_tris;
	const rta_node nl = nodes[L], nr = nodes[R];
	rta_node_pair p;
	p.lo_x[0] = nl.box.min.x; p.lo_x[1] = nr.box.min.x;
	p.lo_y[0] = nl.box.min.y; p.lo_y[1] = nr.box.min.y;
	p.lo_z[0] = nl.box.min.z; p.lo_z[1] = nr.box.min.z;
	p.hi_x[0] = nl.box.max.x; p.hi_x[1] = nr.box.max.x;
	p.hi_y[0] = nl.box.max.y; p.hi_y[1] = nr.box.max.y;
	p.hi_z[0] = nl.box.max.z; p.hi_z[1] = nr.box.max.z;
	p.link[0] = pair_word(nl, L);
	p.link[1] = pair_word(nr, R);
	p.index[0] = L;
	p.index[1] = R;
	pairs[x] = p;
	if (!link_inner(nl.type_left_elems)) mark_leaf_end(tri4, nl);
	if (!link_inner(nr.type_left_elems)) mark_leaf_end(tri4, nr);
	if (!box_fast_ok(nl.box) || !box_fast_ok(nr.box)) atomicOr(flags, 1u);
}

// Batched packing: mesh m's pair records at pairs[node_off], its tri4
// records at tri4[3 * tri_off], its flag word flags[m] (cleared before) and
// its root word (the pair-record word of the root) root_words[m].
__global__ void batch_pack_tris_kernel(const BatchItem *__restrict__ items,
                                       const BatchMesh *__restrict__ meshes,
                                       float4 *__restrict__ tri4,
                                       const rta_triangle *__restrict__ tris) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	const int i = (int)it.block * 256 + (int)threadIdx.x;
	if (i >= (int)m.n) return;
	pack_tri(tri4 + 3 * (size_t)m.tri_off, tris + m.tri_off, i);
}

__global__ void batch_pack_pairs_kernel(const BatchItem *__restrict__ items,
                                        const BatchMesh *__restrict__ meshes,
                                        rta_node_pair *__restrict__ pairs, float4 *__restrict__ tri4,
                                        uint32_t *__restrict__ flags,
                                        const rta_node *__restrict__ nodes,
                                        uint32_t *__restrict__ root_words) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	const int x = (int)it.block * 256 + (int)threadIdx.x;
	if (x >= 2 * (int)m.n - 1) return;
	const rta_node *mn = nodes + m.node_off;
	pack_pair(pairs + m.node_off, tri4 + 3 * (size_t)m.tri_off, flags + it.mesh, mn, x);
	if (x == 0)
		root_words[it.mesh] = link_inner(mn[0].type_left_elems) ? 1u : (mn[0].right_tris << 1);
}

// -------------------------------------------- packed CIS (fast path)

// Blocks are handed out to the 8 XCDs round-robin; mapping block b to ray
// tile (b % 8) * per + b / 8 gives each XCD one contiguous eighth of the
// rays, so neighbouring (coherent) rays share an L2.
constexpr int kXcds = 8;

// Both child boxes of a pair record (q0..q2: lo_x, lo_y, lo_z, hi_x, hi_y,
// hi_z of the left and right child, interleaved).  `fast` rays take the
// branch-free exact-division test, the others the IEEE one.
__device__ inline void pair_boxes(const float4 &q0, const float4 &q1, const float4 &q2, bool fast,
                                  float ox, float oy, float oz, float dx, float dy, float dz,
                                  float yx, float yy, float yz, bool &hl, float &dl, bool &hr,
                                  float &dr) {
	if (fast) {
		hl = box_test_nz(q0.x, q0.z, q1.x, q1.z, q2.x, q2.z, ox, oy, oz, dx, dy, dz, yx, yy, yz, dl);
		hr = box_test_nz(q0.y, q0.w, q1.y, q1.w, q2.y, q2.w, ox, oy, oz, dx, dy, dz, yx, yy, yz, dr);
	} else {
		hl = box_test(q0.x, q0.z, q1.x, q1.z, q2.x, q2.z, ox, oy, oz, dx, dy, dz, dl);
		hr = box_test(q0.y, q0.w, q1.y, q1.w, q2.y, q2.w, ox, oy, oz, dx, dy, dz, dr);
	}
}

// One thread per ray over the child-pair layout.  ANY_HIT selects the
// shadow form (trace_shadow_cis, kernels.cu:857-939): the first triangle a
// leaf accepts ends that leaf and, once the leaf is done, the ray.
template <int BLOCK, bool ANY_HIT>
__global__ __launch_bounds__(BLOCK) void trace_cis_packed_kernel(
    const float4 *__restrict__ tri4, const float4 *__restrict__ pairs,
    const uint32_t *__restrict__ flags, uint32_t root_word, const float *__restrict__ ro,
    const float *__restrict__ rd, const float *__restrict__ rmax, int n_rays,
    const uint32_t *__restrict__ ray_count, rta_hit *__restrict__ hits,
    uint32_t *__restrict__ err) {
	__shared__ uint32_t stack[kStack][BLOCK];
	const int lt = threadIdx.x;
	const int per = gridDim.x / kXcds;
	const int tile = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
	const int tid = tile * BLOCK + lt;
	if (tid >= n_rays) return;
	if (ray_count && (uint32_t)tid >= *ray_count) return;

	const float ox = ro[3 * tid + 0], oy = ro[3 * tid + 1], oz = ro[3 * tid + 2];
	const float dx = rd[3 * tid + 0], dy = rd[3 * tid + 1], dz = rd[3 * tid + 2];
	const float tmax = rmax[tid];
	rta_hit closest = hits[tid];
	closest.t = FLT_MAX;
	const bool fast = ((*flags & 1u) == 0u) && ray_nz_ok(ox, oy, oz, dx, dy, dz);
	const float yx = 1.0f / dx, yy = 1.0f / dy, yz = 1.0f / dz;

	uint32_t curr = root_word;
	int sp = -1;
	for (uint32_t guard = 0;; ++guard) {
		if (guard >= kMaxSteps) { atomicAdd(err, 1u); break; }
		if (curr & 1u) {
			const float4 *p = pairs + 4 * (size_t)(curr >> 1);
			const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
			float dl, dr;
			bool hl, hr;
			pair_boxes(q0, q1, q2, fast, ox, oy, oz, dx, dy, dz, yx, yy, yz, hl, dl, hr, dr);
			hl = hl && dl < closest.t && dl <= tmax;
			hr = hr && dr < closest.t && dr <= tmax;
			const uint32_t wl = __float_as_uint(q3.x), wr = __float_as_uint(q3.y);
			if (hl && hr) {
				if (sp + 1 >= kStack) { atomicAdd(err, 1u); break; }
				if (dl <= dr) { stack[++sp][lt] = wr; curr = wl; }
				else          { stack[++sp][lt] = wl; curr = wr; }
				continue;
			}
			if (hl) { curr = wl; continue; }
			if (hr) { curr = wr; continue; }
		} else {
			uint32_t i = curr >> 1;
			while (true) {
				const float4 v0 = tri4[3 * (size_t)i + 0];
				const float4 v1 = tri4[3 * (size_t)i + 1];
				const float4 v2 = tri4[3 * (size_t)i + 2];
				float t, be, ga;
				if (tri_test(v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z, ox, oy, oz, dx,
				             dy, dz, t, be, ga)) {
					if (t < closest.t && t <= tmax) {
						closest.t = t;
						closest.beta = be;
						closest.gamma = ga;
						closest.ref = i;
						if (ANY_HIT) break;
					}
				}
				if (__float_as_uint(v0.w) != 0u) break;
				++i;
			}
			if (ANY_HIT && closest.t != FLT_MAX) break;
		}
		if (sp < 0) break;
		curr = stack[sp--][lt];
	}
	hits[tid] = closest;
}

// -------------------------------- treelets and seeding from predictions

constexpr int kSeedTop = 4;
constexpr int kMaxTreelets = 1024;

struct CutEntry {
	float lo[3], hi[3];
	uint32_t word, first, size;
};

// Bitonic sort of key/payload pairs over 1024 threads (one element each),
// ascending or descending.  All threads of the block take part.
template <bool DESC>
__device__ inline void block_sort_1024(uint64_t *key, uint32_t *val, int t) {
	for (int k = 2; k <= kMaxTreelets; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1) {
			const int u = t ^ j;
			if (u > t) {
				const bool up = ((t & k) == 0) != DESC;
				const uint64_t a = key[t], b = key[u];
				if ((a > b) == up && a != b) {
					key[t] = b; key[u] = a;
					const uint32_t v = val[t]; val[t] = val[u]; val[u] = v;
				}
			}
			__syncthreads();
		}
}

// The cut in rounds (one block of 1024 threads), an approximation of
// "always split the largest treelet": each round takes the inner treelets
// holding more than half as many triangles as the largest inner one, sorts
// them by (triangle count descending, first ascending), splits as many of
// them as fit under max_treelets into their two children (boxes and words
// from the pair record, ranges from the build), and stops when nothing fits
// or nothing is splittable.  Every round splits at least one treelet.  The
// result is sorted by first position.
__global__ __launch_bounds__(kMaxTreelets) void treelet_cut_kernel(
    const float4 *__restrict__ pairs, uint32_t root_word, const uint2 *__restrict__ ranges,
    uint32_t n_tris, int max_treelets, rta_treelet *__restrict__ table,
    uint32_t *__restrict__ count_out) {
	__shared__ CutEntry ent[kMaxTreelets];
	__shared__ uint64_t key[kMaxTreelets];
	__shared__ uint32_t idx[kMaxTreelets];
	__shared__ int s_count, s_splittable;
	__shared__ uint32_t s_largest;
	const int t = threadIdx.x;
	if (t == 0) {
		ent[0] = CutEntry{{-FLT_MAX, -FLT_MAX, -FLT_MAX}, {FLT_MAX, FLT_MAX, FLT_MAX}, root_word, 0u,
		                  n_tris};
		s_count = 1;
	}
	__syncthreads();
	for (int round = 0; round < kMaxTreelets; ++round) {
		const int K = s_count;
		if (K >= max_treelets) break;
		const bool inner = t < K && (ent[t].word & 1u);
		if (t == 0) {
			s_splittable = 0;
			s_largest = 0u;
		}
		__syncthreads();
		if (inner) atomicMax(&s_largest, ent[t].size);
		__syncthreads();
		uint64_t k = 0;
		if (inner && 2ull * ent[t].size > (uint64_t)s_largest)
			k = ((uint64_t)ent[t].size << 32) | (uint64_t)(0xffffffffu - ent[t].first);
		key[t] = k;
		idx[t] = (uint32_t)t;
		if (k != 0) atomicAdd(&s_splittable, 1);
		__syncthreads();
		block_sort_1024<true>(key, idx, t);
		const int splits = min(s_splittable, max_treelets - K);
		if (splits == 0) break;
		CutEntry a, b;
		int e = -1;
		if (t < splits) {
			e = (int)idx[t];
			const float4 *p = pairs + 4 * (size_t)(ent[e].word >> 1);
			const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
			a = CutEntry{{q0.x, q0.z, q1.x}, {q1.z, q2.x, q2.z}, __float_as_uint(q3.x), 0u, 1u};
			b = CutEntry{{q0.y, q0.w, q1.y}, {q1.w, q2.y, q2.w}, __float_as_uint(q3.y), 0u, 1u};
			auto locate = [&](CutEntry &c) {
				if (c.word & 1u) {
					const uint2 r = ranges[c.word >> 1];
					c.first = r.x;
					c.size = r.y - r.x + 1u;
				} else {
					c.first = c.word >> 1;   // a leaf is never split
				}
			};
			locate(a);
			locate(b);
		}
		__syncthreads();   // every read of ent before the writes below
		if (t < splits) {
			ent[e] = a;
			ent[K + t] = b;
		}
		if (t == 0) s_count = K + splits;
		__syncthreads();
	}
	const int K = s_count;
	key[t] = t < K ? (uint64_t)ent[t].first : ~0ull;
	idx[t] = (uint32_t)t;
	__syncthreads();
	block_sort_1024<false>(key, idx, t);
	if (t < K) {
		const CutEntry &c = ent[idx[t]];
		rta_treelet o;
		for (int j = 0; j < 3; ++j) { o.lo[j] = c.lo[j]; o.hi[j] = c.hi[j]; }
		o.word = c.word;
		o.first = c.first;
		table[t] = o;
	}
	if (t == 0) *count_out = (uint32_t)K;
}

__global__ void treelet_labels_kernel(const rta_hit *__restrict__ hits, int n,
                                      const rta_treelet *__restrict__ table,
                                      const uint32_t *__restrict__ count, int32_t *__restrict__ label,
                                      float *__restrict__ target) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const rta_hit h = hits[i];
	if (h.t == FLT_MAX) {
		label[i] = -1;
		target[i] = 0.0f;
		return;
	}
	// the last treelet whose first position is <= ref (table[0].first = 0)
	uint32_t lo = 0, hi = *count;
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (table[mid].first <= h.ref) lo = mid;
		else hi = mid;
	}
	label[i] = (int32_t)lo;
	target[i] = logf(h.t);
}

// Closest hit of ray tid among the triangles of its kSeedTop predicted
// treelets, each traversed from its root (packed CIS, near child first)
// once the ray meets the root's box before the current seed.  Any triangle
// found is a real hit of the ray, so its t (or FLT_MAX) bounds the ray's
// closest hit from above.
__global__ __launch_bounds__(kTraceBlock) void seed_treelets_kernel(
    const float4 *__restrict__ tri4, const float4 *__restrict__ pairs,
    const uint32_t *__restrict__ flags, const rta_treelet *__restrict__ table,
    const uint32_t *__restrict__ count, const float *__restrict__ ro,
    const float *__restrict__ rd, const float *__restrict__ rmax, int n_rays,
    const int32_t *__restrict__ top, int candidates, float *__restrict__ t_seed,
    uint32_t *__restrict__ err) {
	__shared__ uint32_t stack[kStack][kTraceBlock];
	const int lt = threadIdx.x;
	const int tid = blockIdx.x * kTraceBlock + lt;
	if (tid >= n_rays) return;
	const uint32_t K = *count;
	const float ox = ro[3 * tid + 0], oy = ro[3 * tid + 1], oz = ro[3 * tid + 2];
	const float dx = rd[3 * tid + 0], dy = rd[3 * tid + 1], dz = rd[3 * tid + 2];
	const float tmax = rmax[tid];
	const bool fast = ((*flags & 1u) == 0u) && ray_nz_ok(ox, oy, oz, dx, dy, dz);
	const float yx = 1.0f / dx, yy = 1.0f / dy, yz = 1.0f / dz;
	// the candidates' root boxes first; the treelets are then traversed in
	// order of entry distance, so an early hit prunes the later ones.  A box
	// entry distance is finite, so +inf marks a candidate that is skipped.
	float dk[kSeedTop];
	uint32_t wk[kSeedTop];
	for (int k = 0; k < kSeedTop; ++k) {
		const uint32_t c = (uint32_t)top[(size_t)tid * kSeedTop + k];
		dk[k] = __builtin_inff();
		wk[k] = 0u;
		if (k >= candidates || c >= K) continue;
		const rta_treelet tr = table[c];
		float dist;
		if (box_test(tr.lo[0], tr.lo[1], tr.lo[2], tr.hi[0], tr.hi[1], tr.hi[2], ox, oy, oz, dx, dy,
		             dz, dist) &&
		    dist <= tmax) {
			dk[k] = dist;
			wk[k] = tr.word;
		}
	}
	auto order = [&](int a, int b) {
		if (dk[b] < dk[a]) {
			const float d = dk[a]; dk[a] = dk[b]; dk[b] = d;
			const uint32_t w = wk[a]; wk[a] = wk[b]; wk[b] = w;
		}
	};
	order(0, 1); order(2, 3); order(0, 2); order(1, 3); order(1, 2);
	float ct = FLT_MAX;
	uint32_t guard = 0;
	for (int k = 0; k < kSeedTop; ++k) {
		if (!(dk[k] < ct)) continue;     // skipped, or entered after the seed
		uint32_t curr = wk[k];
		int sp = -1;
		for (;; ++guard) {
			if (guard >= kMaxSteps) { atomicAdd(err, 1u); break; }
			if (curr & 1u) {
				const float4 *p = pairs + 4 * (size_t)(curr >> 1);
				const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
				float dl, dr;
				bool hl, hr;
				pair_boxes(q0, q1, q2, fast, ox, oy, oz, dx, dy, dz, yx, yy, yz, hl, dl, hr, dr);
				const uint32_t wl = __float_as_uint(q3.x), wr = __float_as_uint(q3.y);
				hl = hl && dl < ct && dl <= tmax;
				hr = hr && dr < ct && dr <= tmax;
				if (hl && hr) {
					if (sp + 1 >= kStack) { atomicAdd(err, 1u); break; }
					if (dl <= dr) { stack[++sp][lt] = wr; curr = wl; }
					else          { stack[++sp][lt] = wl; curr = wr; }
					continue;
				}
				if (hl) { curr = wl; continue; }
				if (hr) { curr = wr; continue; }
			} else {
				uint32_t i = curr >> 1;
				while (true) {
					const float4 v0 = tri4[3 * (size_t)i + 0];
					const float4 v1 = tri4[3 * (size_t)i + 1];
					const float4 v2 = tri4[3 * (size_t)i + 2];
					float t, be, ga;
					if (tri_test(v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z, ox, oy, oz, dx,
					             dy, dz, t, be, ga) &&
					    t < ct && t <= tmax)
						ct = t;
					if (__float_as_uint(v0.w) != 0u) break;
					++i;
				}
			}
			if (sp < 0) break;
			curr = stack[sp--][lt];
		}
	}
	t_seed[tid] = ct;
}

__global__ void seed_max_t_kernel(float *__restrict__ out, const float *__restrict__ max_t,
                                  const float *__restrict__ t_seed, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = fminf(max_t[i], seed_bound(t_seed[i]));
}

// ---------------------------- persistent packed CIS with ray refill

__global__ void fill_queue_kernel(uint32_t *queue) {
	if (threadIdx.x == 0) *queue = 0u;
}

// Persistent waves: each lane traces one ray at a time; when at least
// `refill_min` lanes of a wave are idle the wave takes that many new rays
// (consecutive ids, so they are coherent) from a global queue counter with
// one atomic.  A wave never waits for another, and it leaves the loop once
// the queue is exhausted and all of its lanes are idle.
//
// Each traversal step of the wave is either an inner-node step (every lane
// standing at an inner node visits it) or a leaf step (every lane standing
// at a leaf tests its triangles and pops): leaf steps run only once at least
// `leaf_min` lanes wait at a leaf or no lane stands at an inner node, so the
// wave does not pay for both code paths in most steps.  Waiting changes
// when a lane advances, never what it does next, so every ray runs the
// identical traversal and results equal trace_cis_packed_kernel's; a miss
// stores only t = FLT_MAX over the reset sentinel (the reference writes the
// sentinel back with t = FLT_MAX, kernels.cu:793-794,853).  ANY_HIT as in
// trace_cis_packed_kernel.
//
// BATCHED: rays of many meshes in one launch.  Ray r belongs to the mesh
// whose [ray_off, ray_off + n_rays) holds it; that mesh's pair records start
// at pairs[node_off], its tri4 records at tri4[3 * tri_off], its flag word
// is flags[mesh] and its root word root_words[mesh]; hit refs are local to
// the mesh.
//
// SEEDED (closest hit, one mesh): each ray first traverses the treelet
// seed.top[4 * ray] proposes (seed.table entry, when its root box is met
// within max_t) until the first leaf that holds a hit, then the whole tree
// with max_t' = min(max_t, the smallest t of that phase) and its closest hit
// reset to FLT_MAX.  Any hit of the ray bounds its closest one, so the
// records are the plain traversal's.  Only the second phase writes the
// record.
struct SeedArgs {
	const rta_treelet *table = nullptr;
	const uint32_t *count = nullptr;
	const int32_t *top = nullptr;
};

// STACK: the LDS stack's entries (kStack, or fewer for a smaller LDS
// footprint and so more resident waves).  With STACK < kStack a ray that
// would push beyond it is abandoned and appended to the overflow list ovf
// (ovf[0] the count, rays from ovf[1]); the caller traces those rays again
// with kStack entries.  ray_list (optional): trace the rays ray_list[1 ..
// ray_list[0]] instead of 0 .. n-1.
template <int BLOCK, bool ANY_HIT, bool BATCHED, bool SEEDED = false, int STACK = kStack>
__global__ __launch_bounds__(BLOCK) void trace_cis_persistent_kernel(
    const float4 *__restrict__ tri4, const float4 *__restrict__ pairs,
    const uint32_t *__restrict__ flags, uint32_t root_word, const float *__restrict__ ro,
    const float *__restrict__ rd, const float *__restrict__ rmax, int n_rays,
    const uint32_t *__restrict__ ray_count, rta_hit *__restrict__ hits,
    uint32_t *__restrict__ err, uint32_t *__restrict__ queue, int refill_min, int steps,
    int leaf_min, const BatchMesh *__restrict__ meshes, int n_meshes,
    const uint32_t *__restrict__ root_words, uint32_t *__restrict__ ovf,
    const uint32_t *__restrict__ ray_list, SeedArgs seed) {
	static_assert(!SEEDED || (!ANY_HIT && !BATCHED), "seeding: closest hit on one mesh");
	static_assert(STACK <= kStack, "at most kStack entries");
	__shared__ uint32_t stack[STACK][BLOCK];
	const int lt = threadIdx.x;
	const int lane = __lane_id();
	const uint64_t lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	uint32_t n = (uint32_t)n_rays;
	if (ray_count) n = min(n, *ray_count);
	if (ray_list) n = min(n, ray_list[0]);
	// nothing to trace (an empty overflow list, most often): leave before
	// the queue atomics of every wave in the grid
	if (n == 0u) return;
	const bool scene_fast = BATCHED ? false : (*flags & 1u) == 0u;
	const uint32_t n_treelets = SEEDED ? *seed.count : 0u;
	bool seeding = false;   // SEEDED: the lane is in its treelet phase

	bool active = false, exhausted = false;
	uint32_t ray = 0, curr = 0;
	uint32_t node_base = 0, tri_base = 0;
	int sp = -1;
	float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, yx = 0, yy = 0, yz = 0, tmax = 0;
	float ct = FLT_MAX, cb = 0, cg = 0;
	uint32_t cref = 0;
	bool fast = false;
	// Rounds (of at most `steps` wave-steps) since the last refill.  Every
	// wave-step advances at least one lane and each lane holds at most one
	// ray per refill, so with a valid tree (at most 2 * nodes < kMaxSteps
	// steps per ray) a wave takes fewer than 64 * kMaxSteps wave-steps, and
	// so at most round_limit rounds, between refills; a malformed tree ends
	// in an error count per active ray instead of an endless loop.
	uint64_t rounds = 0;
	const uint64_t round_limit = 64ull * kMaxSteps / (uint64_t)steps + 2;
	// write the ray's record (hits[ray].t = FLT_MAX alone on a miss) and
	// free the lane
	auto finish = [&]() {
		if (ct != FLT_MAX) {
			rta_hit hres;
			hres.t = ct; hres.beta = cb; hres.gamma = cg; hres.ref = cref;
			hits[ray] = hres;
		} else {
			hits[ray].t = FLT_MAX;
		}
		active = false;
	};

	while (true) {
		const uint64_t idle = __ballot(!active);
		const int nidle = __popcll(idle);
		if (!exhausted && nidle >= refill_min) {
			rounds = 0;
			const int leader = __ffsll((unsigned long long)idle) - 1;
			uint32_t base = 0;
			if (lane == leader) base = atomicAdd(queue, (uint32_t)nidle);
			base = __shfl(base, leader);
			if (base + (uint32_t)nidle >= n) exhausted = true;
			if (!active) {
				const uint32_t my = base + (uint32_t)__popcll(idle & lower);
				if (my < n) {
					ray = ray_list ? ray_list[1 + my] : my;
					ox = ro[3 * (size_t)ray + 0]; oy = ro[3 * (size_t)ray + 1]; oz = ro[3 * (size_t)ray + 2];
					dx = rd[3 * (size_t)ray + 0]; dy = rd[3 * (size_t)ray + 1]; dz = rd[3 * (size_t)ray + 2];
					tmax = rmax[ray];
					if (BATCHED) {
						// last mesh whose first ray is <= ray
						int lo = 0, hi = n_meshes;
						while (hi - lo > 1) {
							const int mid = (lo + hi) >> 1;
							if (meshes[mid].ray_off <= ray) lo = mid;
							else hi = mid;
						}
						node_base = meshes[lo].node_off;
						tri_base = meshes[lo].tri_off;
						curr = root_words[lo];
						fast = (flags[lo] & 1u) == 0u && ray_nz_ok(ox, oy, oz, dx, dy, dz);
					} else {
						curr = root_word;
						fast = scene_fast && ray_nz_ok(ox, oy, oz, dx, dy, dz);
					}
					if (SEEDED) {
						seeding = false;
						const uint32_t c = (uint32_t)seed.top[4 * (size_t)ray];
						if (c < n_treelets) {
							const rta_treelet tr = seed.table[c];
							float dist;
							if (box_test(tr.lo[0], tr.lo[1], tr.lo[2], tr.hi[0], tr.hi[1], tr.hi[2], ox,
							             oy, oz, dx, dy, dz, dist) &&
							    dist <= tmax) {
								curr = tr.word;
								seeding = true;
							}
						}
					}
					yx = 1.0f / dx; yy = 1.0f / dy; yz = 1.0f / dz;
					ct = FLT_MAX;
					sp = -1;
					active = true;
				}
			}
		}
		// with refill_min <= 64 an all-idle wave has just refilled, so an
		// all-idle wave here means the queue is exhausted
		if (__ballot(active) == 0ull) break;
		if (++rounds > round_limit) {
			// malformed tree: every active ray ends with an error count
			if (active) {
				atomicAdd(err, 1u);
				finish();
			}
			continue;
		}
		for (int k = 0; k < steps; ++k) {
			const bool is_inner = (curr & 1u) != 0u;
			const uint64_t act = __ballot(active);
			if (act == 0ull) break;
			const uint64_t inner_lanes = __ballot(active && is_inner);
			// Wave-uniform choice of this step's kind: a leaf step once
			// leaf_min of 64 lanes, scaled to the lanes still active, wait at
			// leaves (a draining wave does not hold its last rays there).
			const int n_leaf = __popcll(act & ~inner_lanes);
			const int thr = (__popcll(act) * leaf_min + 63) >> 6;
			const bool leaf_step = inner_lanes == 0ull || n_leaf >= thr;
			if (!active || is_inner == leaf_step) continue;
			bool done, overflow = false;
			if (!leaf_step) {
				const float4 *p = pairs + 4 * ((size_t)node_base + (curr >> 1));
				const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
				float dl, dr;
				bool hl, hr;
				pair_boxes(q0, q1, q2, fast, ox, oy, oz, dx, dy, dz, yx, yy, yz, hl, dl, hr, dr);
				hl = hl && dl < ct && dl <= tmax;
				hr = hr && dr < ct && dr <= tmax;
				const uint32_t wl = __float_as_uint(q3.x), wr = __float_as_uint(q3.y);
				// The reference's decision (kernels.cu:818-835) without
				// branches: both kept -> enter the nearer, the left one on a
				// tie, and push the other; one kept -> enter it; none -> pop,
				// or finish on an empty stack.  The far child is written one
				// slot above the top on every step (after the top is read);
				// only a push makes that slot part of the stack.  On a full
				// stack (STACK entries) there is no slot above the top: the
				// top entry is written back with its own value, so a step that
				// keeps one child or none leaves the stack intact, and a push
				// ends the ray with an error count (STACK = kStack) or on the
				// overflow list (STACK < kStack).
				const bool both = hl && hr, any = hl || hr;
				const bool enter_left = both ? (dl <= dr) : hl;
				const bool full = sp + 1 >= STACK;
				const uint32_t popped = stack[sp >= 0 ? sp : 0][lt];
				stack[full ? STACK - 1 : sp + 1][lt] = full ? popped : (enter_left ? wr : wl);
				curr = any ? (enter_left ? wl : wr) : popped;
				overflow = both && full;
				done = overflow || (!any && sp < 0);
				sp += both ? 1 : (any ? 0 : -1);
			} else {
				uint32_t i = curr >> 1;
				const float4 *tb = tri4 + 3 * (size_t)tri_base;
				while (true) {
					const float4 v0 = tb[3 * (size_t)i + 0];
					const float4 v1 = tb[3 * (size_t)i + 1];
					const float4 v2 = tb[3 * (size_t)i + 2];
					float t, be, ga;
					if (tri_test(v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z, ox, oy, oz,
					             dx, dy, dz, t, be, ga)) {
						if (t < ct && t <= tmax) {
							ct = t; cb = be; cg = ga; cref = i;
							if (ANY_HIT) break;
						}
					}
					if (__float_as_uint(v0.w) != 0u) break;
					++i;
				}
				// pop, or finish (empty stack, a shadow ray's first hit, or the
				// end of a treelet phase at its first hit leaf: any real hit
				// is a valid bound)
				curr = stack[sp >= 0 ? sp : 0][lt];
				done = (ANY_HIT && ct != FLT_MAX) || (SEEDED && seeding && ct != FLT_MAX) || sp < 0;
				--sp;
			}
			if (done) {
				if (overflow) {
					if (STACK < kStack) ovf[1 + atomicAdd(ovf, 1u)] = ray;   // traced again
					else atomicAdd(err, 1u);
				}
				if (SEEDED && seeding) {
					// the treelet phase ends: its closest t (a real hit, or
					// FLT_MAX), widened by seed_bound's margin, bounds the
					// traversal of the whole tree
					tmax = fminf(tmax, seed_bound(ct));
					ct = FLT_MAX;
					curr = root_word;
					sp = -1;
					seeding = false;
				} else {
					finish();
				}
			}
		}
	}
}

// ------------------------------------------------------- bouncing

constexpr int kBounceBlock = 256;

__global__ __launch_bounds__(kBounceBlock) void bounce_count_kernel(const rta_hit *__restrict__ hits,
                                                                    int n, uint32_t *__restrict__ block_counts) {
	__shared__ uint32_t wsum[kBounceBlock / kWave];
	const int i = blockIdx.x * kBounceBlock + threadIdx.x;
	const bool hit = i < n && hits[i].t != FLT_MAX;
	const uint64_t b = __ballot(hit);
	if (__lane_id() == 0) wsum[threadIdx.x / kWave] = (uint32_t)__popcll(b);
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t s = 0;
		for (int w = 0; w < kBounceBlock / kWave; ++w) s += wsum[w];
		block_counts[blockIdx.x] = s;
	}
}

// single block: exclusive scan of block counts, total into *count
__global__ __launch_bounds__(1024) void bounce_scan_kernel(uint32_t *__restrict__ block_counts, int nb,
                                                           uint32_t *__restrict__ count) {
	__shared__ uint32_t wsum[1024 / kWave];
	__shared__ uint32_t carry_s;
	if (threadIdx.x == 0) carry_s = 0;
	__syncthreads();
	for (int base = 0; base < nb; base += 1024) {
		const int idx = base + threadIdx.x;
		const uint32_t v = idx < nb ? block_counts[idx] : 0u;
		uint32_t x = v;
		const int lane = __lane_id();
		for (int off = 1; off < kWave; off <<= 1) {
			uint32_t y = __shfl_up(x, off);
			if (lane >= off) x += y;
		}
		if (lane == kWave - 1) wsum[threadIdx.x / kWave] = x;
		__syncthreads();
		uint32_t before = 0, all = 0;
		for (int w = 0; w < 1024 / kWave; ++w) {
			const uint32_t s = wsum[w];
			if (w < (int)(threadIdx.x / kWave)) before += s;
			all += s;
		}
		const uint32_t carry = carry_s;
		if (idx < nb) block_counts[idx] = carry + before + x - v;
		__syncthreads();
		if (threadIdx.x == 0) carry_s = carry + all;
		__syncthreads();
	}
	if (threadIdx.x == 0) *count = carry_s;
}

__device__ inline void bounce_ray(const rta_hit &h, const rta_triangle *tris, float ox, float oy,
                                  float oz, float dx, float dy, float dz, uint64_t seed, uint32_t i,
                                  float *o_out, float *d_out) {
	const float *f = reinterpret_cast<const float *>(tris + h.ref);
	const float e1x = f[3] - f[0], e1y = f[4] - f[1], e1z = f[5] - f[2];
	const float e2x = f[6] - f[0], e2y = f[7] - f[1], e2z = f[8] - f[2];
	float nx = e1y * e2z - e1z * e2y;
	float ny = e1z * e2x - e1x * e2z;
	float nz = e1x * e2y - e1y * e2x;
	const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
	nx = nx / nl; ny = ny / nl; nz = nz / nl;
	if (nx * dx + ny * dy + nz * dz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
	const float px = ox + h.t * dx, py = oy + h.t * dy, pz = oz + h.t * dz;
	o_out[0] = px + 1e-4f * nx;
	o_out[1] = py + 1e-4f * ny;
	o_out[2] = pz + 1e-4f * nz;
	// uniform point on the unit sphere, Marsaglia (1972)
	const uint64_t base = mix64(seed ^ mix64((uint64_t)i));
	float x1 = 0.0f, x2 = 0.0f, s = 0.0f;
	for (uint32_t it = 0; it < 64; ++it) {
		const uint64_t r = mix64(base + it);
		const float u1 = (float)(uint32_t)(r >> 40) * (1.0f / 16777216.0f);
		const float u2 = (float)(uint32_t)((r >> 16) & 0xffffffu) * (1.0f / 16777216.0f);
		const float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
		const float q = a * a + b * b;
		if (q < 1.0f) { x1 = a; x2 = b; s = q; break; }
	}
	const float sq = sqrtf(1.0f - s);
	const float sx = (2.0f * x1) * sq, sy = (2.0f * x2) * sq, sz = 1.0f - 2.0f * s;
	float vx = nx + sx, vy = ny + sy, vz = nz + sz;
	const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
	if (vl > 0.0f) { vx = vx / vl; vy = vy / vl; vz = vz / vl; }
	else { vx = nx; vy = ny; vz = nz; }
	d_out[0] = vx; d_out[1] = vy; d_out[2] = vz;
}

// Shadow ray of a hit towards a point light: origin as for the bounce
// (hit point + 1e-4 * normal facing the incoming ray), direction the unit
// vector to the light, max_t the distance to it.  A light at the origin
// itself gives direction n and max_t 0 (never occluded).
__device__ inline void shadow_ray(const rta_hit &h, const rta_triangle *tris, float ox, float oy,
                                  float oz, float dx, float dy, float dz, float lx, float ly,
                                  float lz, float *o_out, float *d_out, float *t_out) {
	const float *f = reinterpret_cast<const float *>(tris + h.ref);
	const float e1x = f[3] - f[0], e1y = f[4] - f[1], e1z = f[5] - f[2];
	const float e2x = f[6] - f[0], e2y = f[7] - f[1], e2z = f[8] - f[2];
	float nx = e1y * e2z - e1z * e2y;
	float ny = e1z * e2x - e1x * e2z;
	float nz = e1x * e2y - e1y * e2x;
	const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
	nx = nx / nl; ny = ny / nl; nz = nz / nl;
	if (nx * dx + ny * dy + nz * dz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
	const float px = ox + h.t * dx, py = oy + h.t * dy, pz = oz + h.t * dz;
	const float sx = px + 1e-4f * nx, sy = py + 1e-4f * ny, sz = pz + 1e-4f * nz;
	o_out[0] = sx; o_out[1] = sy; o_out[2] = sz;
	float vx = lx - sx, vy = ly - sy, vz = lz - sz;
	const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
	if (vl > 0.0f) {
		d_out[0] = vx / vl; d_out[1] = vy / vl; d_out[2] = vz / vl;
		*t_out = vl;
	} else {
		d_out[0] = nx; d_out[1] = ny; d_out[2] = nz;
		*t_out = 0.0f;
	}
}

// Writes one derived ray per hit, compacted in ray order.  SHADOW selects
// shadow rays (light at lx, ly, lz) over diffuse bounces (seed).
template <bool SHADOW>
__global__ __launch_bounds__(kBounceBlock) void bounce_emit_kernel(
    float *__restrict__ out_o, float *__restrict__ out_d, float *__restrict__ out_t,
    uint32_t *__restrict__ src_index, const uint32_t *__restrict__ block_offsets,
    const rta_hit *__restrict__ hits, const rta_triangle *__restrict__ tris,
    const float *__restrict__ ro, const float *__restrict__ rd, int n, uint64_t seed, float lx,
    float ly, float lz) {
	__shared__ uint32_t wsum[kBounceBlock / kWave];
	const int i = blockIdx.x * kBounceBlock + threadIdx.x;
	rta_hit h;
	bool hit = false;
	if (i < n) { h = hits[i]; hit = h.t != FLT_MAX; }
	const uint64_t b = __ballot(hit);
	const int lane = __lane_id(), wv = threadIdx.x / kWave;
	if (lane == 0) wsum[wv] = (uint32_t)__popcll(b);
	__syncthreads();
	if (!hit) return;
	uint32_t before = 0;
	for (int w = 0; w < wv; ++w) before += wsum[w];
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	const uint32_t k = block_offsets[blockIdx.x] + before + (uint32_t)__popcll(b & lt);
	float o[3], d[3], t = FLT_MAX;
	if (SHADOW)
		shadow_ray(h, tris, ro[3 * i], ro[3 * i + 1], ro[3 * i + 2], rd[3 * i], rd[3 * i + 1],
		           rd[3 * i + 2], lx, ly, lz, o, d, &t);
	else
		bounce_ray(h, tris, ro[3 * i], ro[3 * i + 1], ro[3 * i + 2], rd[3 * i], rd[3 * i + 1],
		           rd[3 * i + 2], seed, (uint32_t)i, o, d);
	out_o[3 * (size_t)k + 0] = o[0]; out_o[3 * (size_t)k + 1] = o[1]; out_o[3 * (size_t)k + 2] = o[2];
	out_d[3 * (size_t)k + 0] = d[0]; out_d[3 * (size_t)k + 1] = d[1]; out_d[3 * (size_t)k + 2] = d[2];
	out_t[k] = t;
	src_index[k] = (uint32_t)i;
}

}  // namespace rta_hip

namespace rta_hip {

// Grid of the persistent kernel: as many blocks as fit on every CU (or
// blocks_per_cu), never more than the rays need.  The CU count and the
// occupancy are queried once per device and kernel instance.  With STACK <
// kStack the rays that overflow the stack (the list ovf, n_rays + 1 words)
// are traced again by a kStack launch over that list.
template <int BLOCK, bool ANY_HIT, bool BATCHED, bool SEEDED = false, int STACK = kStack>
int launch_persistent(const float *tri4, const rta_node_pair *pairs, const uint32_t *flags,
                      uint32_t root_word, const float *ro, const float *rd, const float *rmax,
                      int n_rays, const uint32_t *ray_count, rta_hit *hits, uint32_t *err,
                      uint32_t *queue, int refill_min, int steps, int blocks_per_cu,
                      int leaf_min, const BatchMesh *meshes, int n_meshes,
                      const uint32_t *root_words, hipStream_t st, SeedArgs seed = SeedArgs{},
                      uint32_t *ovf = nullptr, const uint32_t *ray_list = nullptr) {
	constexpr int kMaxDevices = 64;
	static int cus[kMaxDevices] = {}, occ[kMaxDevices] = {};
	int dev = 0;
	RTA_HIP_TRY(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices) return invalid("launch_persistent: device id out of range");
	if (STACK < kStack && !ovf) return invalid("launch_persistent: a short stack needs an overflow list");
	if (cus[dev] == 0) {
		int c = 0, o = 0;
		RTA_HIP_TRY(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
		RTA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
		    &o, trace_cis_persistent_kernel<BLOCK, ANY_HIT, BATCHED, SEEDED, STACK>, BLOCK, 0));
		occ[dev] = o < 1 ? 1 : o;
		cus[dev] = c;
	}
	const int bpc = blocks_per_cu > 0 ? blocks_per_cu : occ[dev];
	const long long want = (long long)cus[dev] * bpc;
	const long long need = ((long long)n_rays + BLOCK - 1) / BLOCK;
	const unsigned blocks = (unsigned)(want < need ? want : need);
	if (STACK < kStack) {
		const int rc = rta_hip_fill_u32(ovf, 0u, 1, reinterpret_cast<rta_stream>(st));
		if (rc != RTA_OK) return rc;
	}
	hipLaunchKernelGGL(fill_queue_kernel, dim3(1), dim3(64), 0, st, queue);
	RTA_LAUNCH_CHECK("fill_queue_kernel");
	hipLaunchKernelGGL((trace_cis_persistent_kernel<BLOCK, ANY_HIT, BATCHED, SEEDED, STACK>),
	                   dim3(blocks), dim3(BLOCK), 0, st, (const float4 *)tri4,
	                   (const float4 *)pairs, flags, root_word, ro, rd, rmax, n_rays, ray_count,
	                   hits, err, queue, refill_min, steps, leaf_min, meshes, n_meshes,
	                   root_words, STACK < kStack ? ovf : nullptr, ray_list, seed);
	RTA_LAUNCH_CHECK("trace_cis_persistent_kernel");
	if (STACK < kStack)
		return launch_persistent<BLOCK, ANY_HIT, BATCHED, SEEDED, kStack>(
		    tri4, pairs, flags, root_word, ro, rd, rmax, n_rays, nullptr, hits, err, queue,
		    refill_min, steps, blocks_per_cu, leaf_min, meshes, n_meshes, root_words, st, seed,
		    nullptr, ovf);
	return RTA_OK;
}

// Validated launch shape with defaults filled in.
static int resolve_config(const char *name, const rta_trace_config *cfg, rta_trace_config &c) {
	c = {64, 16, 32, 0, 0, 0};
	if (cfg) c = *cfg;
	if (c.leaf_min == 0) c.leaf_min = 10;
	if (c.stack == 0) c.stack = kStack;
	if (c.refill_min < 1 || c.refill_min > 64 || c.steps < 1 || c.blocks_per_cu < 0 ||
	    c.leaf_min < 1 || c.leaf_min > 64 || (c.block != 64 && c.block != 128 && c.block != 256) ||
	    (c.stack != 16 && c.stack != 24 && c.stack != kStack)) {
		set_error(std::string(name) + ": bad config (block 64/128/256, refill_min and leaf_min "
		                              "1..64, steps >= 1, blocks_per_cu >= 0, stack 16/24/32)");
		return RTA_E_INVALID;
	}
	return RTA_OK;
}

// stack 16 and 24 need an overflow list (n_rays + 1 words)
static int require_full_stack(const char *name, const rta_trace_config &c,
                              const uint32_t *ovf = nullptr) {
	if (c.stack == kStack || ovf) return RTA_OK;
	set_error(std::string(name) + ": stack 16 and 24 need an overflow list (the _ovf entry "
	                              "points, or the batch API)");
	return RTA_E_INVALID;
}

template <int BLOCK, bool ANY_HIT, bool BATCHED>
int launch_stack(const rta_trace_config &c, const float *tri4, const rta_node_pair *pairs,
                 const uint32_t *flags, uint32_t root_word, const float *ro, const float *rd,
                 const float *rmax, int n_rays, const uint32_t *ray_count, rta_hit *hits,
                 uint32_t *err, uint32_t *queue, const BatchMesh *meshes, int n_meshes,
                 const uint32_t *root_words, uint32_t *ovf, hipStream_t st) {
	if (c.stack == 16)
		return launch_persistent<BLOCK, ANY_HIT, BATCHED, false, 16>(
		    tri4, pairs, flags, root_word, ro, rd, rmax, n_rays, ray_count, hits, err, queue,
		    c.refill_min, c.steps, c.blocks_per_cu, c.leaf_min, meshes, n_meshes, root_words, st,
		    SeedArgs{}, ovf);
	if (c.stack == 24)
		return launch_persistent<BLOCK, ANY_HIT, BATCHED, false, 24>(
		    tri4, pairs, flags, root_word, ro, rd, rmax, n_rays, ray_count, hits, err, queue,
		    c.refill_min, c.steps, c.blocks_per_cu, c.leaf_min, meshes, n_meshes, root_words, st,
		    SeedArgs{}, ovf);
	return launch_persistent<BLOCK, ANY_HIT, BATCHED>(
	    tri4, pairs, flags, root_word, ro, rd, rmax, n_rays, ray_count, hits, err, queue,
	    c.refill_min, c.steps, c.blocks_per_cu, c.leaf_min, meshes, n_meshes, root_words, st);
}

template <bool ANY_HIT, bool BATCHED>
int launch_persistent_cfg(const rta_trace_config &c, const float *tri4, const rta_node_pair *pairs,
                          const uint32_t *flags, uint32_t root_word, const float *ro,
                          const float *rd, const float *rmax, int n_rays,
                          const uint32_t *ray_count, rta_hit *hits, uint32_t *err,
                          uint32_t *queue, const BatchMesh *meshes, int n_meshes,
                          const uint32_t *root_words, hipStream_t st, uint32_t *ovf = nullptr) {
	switch (c.block) {
	case 64:
		return launch_stack<64, ANY_HIT, BATCHED>(c, tri4, pairs, flags, root_word, ro, rd, rmax,
		                                          n_rays, ray_count, hits, err, queue, meshes,
		                                          n_meshes, root_words, ovf, st);
	case 128:
		return launch_stack<128, ANY_HIT, BATCHED>(c, tri4, pairs, flags, root_word, ro, rd, rmax,
		                                           n_rays, ray_count, hits, err, queue, meshes,
		                                           n_meshes, root_words, ovf, st);
	default:
		return launch_stack<256, ANY_HIT, BATCHED>(c, tri4, pairs, flags, root_word, ro, rd, rmax,
		                                           n_rays, ray_count, hits, err, queue, meshes,
		                                           n_meshes, root_words, ovf, st);
	}
}

template <bool ANY_HIT>
int trace_packed(const char *name, const float *tri4, const rta_node_pair *pairs,
                 const uint32_t *flags, uint32_t root_link, uint32_t root_index,
                 const float *ray_orig, const float *ray_dir, const float *max_t, int n_rays,
                 const uint32_t *ray_count, rta_hit *hits, uint32_t *err, hipStream_t st) {
	if (!tri4 || !pairs || !flags || !ray_orig || !ray_dir || !max_t || !hits || !err ||
	    n_rays <= 0) {
		set_error(std::string(name) + ": bad arguments");
		return RTA_E_INVALID;
	}
	// the root as its parent's pair record would describe it
	const uint32_t root_word = link_inner(root_link) ? 1u : (root_index << 1);
	const unsigned tiles = grid_for(n_rays, kTraceBlock);
	const unsigned blocks = (tiles + kXcds - 1) / kXcds * kXcds;
	hipLaunchKernelGGL((trace_cis_packed_kernel<kTraceBlock, ANY_HIT>), dim3(blocks),
	                   dim3(kTraceBlock), 0, st, (const float4 *)tri4, (const float4 *)pairs, flags,
	                   root_word, ray_orig, ray_dir, max_t, n_rays, ray_count, hits, err);
	RTA_LAUNCH_CHECK("trace_cis_packed_kernel");
	return RTA_OK;
}

template <bool ANY_HIT>
int trace_persistent(const char *name, const float *tri4, const rta_node_pair *pairs,
                     const uint32_t *flags, uint32_t root_link, uint32_t root_index,
                     const float *ray_orig, const float *ray_dir, const float *max_t, int n_rays,
                     const uint32_t *ray_count, rta_hit *hits, uint32_t *err, uint32_t *queue,
                     const rta_trace_config *cfg, hipStream_t st, uint32_t *ovf = nullptr) {
	const ScopedRange range(name);
	if (!tri4 || !pairs || !flags || !ray_orig || !ray_dir || !max_t || !hits || !err ||
	    !queue || n_rays <= 0) {
		set_error(std::string(name) + ": bad arguments");
		return RTA_E_INVALID;
	}
	rta_trace_config c;
	int rc = resolve_config(name, cfg, c);
	if (rc == RTA_OK) rc = require_full_stack(name, c, ovf);
	if (rc != RTA_OK) return rc;
	const uint32_t root_word = link_inner(root_link) ? 1u : (root_index << 1);
	return launch_persistent_cfg<ANY_HIT, false>(c, tri4, pairs, flags, root_word, ray_orig,
	                                             ray_dir, max_t, n_rays, ray_count, hits, err,
	                                             queue, nullptr, 0, nullptr, st,
	                                             c.stack < kStack ? ovf : nullptr);
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" {

int rta_reset_intersections(rta_hit *hits, int w, int h, rta_stream s) {
	if (w <= 0 || h <= 0 || !hits) return invalid("rta_reset_intersections: bad arguments");
	const long long n = (long long)w * h;
	hipLaunchKernelGGL(reset_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(s), hits, (int)n);
	RTA_LAUNCH_CHECK("reset_kernel");
	return RTA_OK;
}

// camera basis exactly as cam_ray_generator_shirley::generate_ray_dir
// (librta/raytrav.h:308-333): W = dir/|dir|, U = (up x W)/|up x W|, V = W x U
static CamBasis make_basis(float fovy, float aspect, const float *vd, const float *pos,
                           const float *up) {
	CamBasis cb;
	float fh = (float)((double)fovy / 2.0);
	cb.height = tanf((float)(M_PI * (double)fh / 180.0f));
	cb.width = aspect * cb.height;
	const float wl = sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
	cb.wx = vd[0] / wl; cb.wy = vd[1] / wl; cb.wz = vd[2] / wl;
	float tx = up[1] * cb.wz - up[2] * cb.wy;
	float ty = up[2] * cb.wx - up[0] * cb.wz;
	float tz = up[0] * cb.wy - up[1] * cb.wx;
	const float tl = sqrtf(tx * tx + ty * ty + tz * tz);
	cb.ux = tx / tl; cb.uy = ty / tl; cb.uz = tz / tl;
	cb.vx = cb.wy * cb.uz - cb.wz * cb.uy;
	cb.vy = cb.wz * cb.ux - cb.wx * cb.uz;
	cb.vz = cb.wx * cb.uy - cb.wy * cb.ux;
	cb.px = pos[0]; cb.py = pos[1]; cb.pz = pos[2];
	return cb;
}

int rta_setup_shirley_rays(float *dirs, float *origins, float *max_t, float fovy, float aspect,
                           int w, int h, const float *view_dir, const float *pos, const float *up,
                           float maxt, rta_stream s) {
	if (!dirs || !origins || !max_t || !view_dir || !pos || !up || w <= 0 || h <= 0)
		return invalid("rta_setup_shirley_rays: bad arguments");
	const CamBasis cb = make_basis(fovy, aspect, view_dir, pos, up);
	dim3 threads(16, 16);
	dim3 blocks(grid_for(w, 16), grid_for(h, 16));
	hipLaunchKernelGGL(shirley_kernel, blocks, threads, 0, as_stream(s), dirs, origins, max_t, w, h,
	                   cb, maxt);
	RTA_LAUNCH_CHECK("shirley_kernel");
	return RTA_OK;
}

int rta_trace_cis(const rta_triangle *tris, int n_tris, const rta_node *nodes,
                  const float *ray_orig, const float *ray_dir, const float *max_t, int w, int h,
                  rta_hit *hits, uint32_t *stats, uint32_t *err, rta_stream s) {
	if (!tris || n_tris <= 0 || !nodes || !ray_orig || !ray_dir || !max_t || !hits || !err ||
	    w <= 0 || h <= 0)
		return invalid("rta_trace_cis: bad arguments");
	const long long n = (long long)w * h;
	if (n > 0x7fffffffLL) return invalid("rta_trace_cis: too many rays");
	if (stats)
		hipLaunchKernelGGL(trace_cis_kernel<true>, dim3(grid_for(n, kTraceBlock)), dim3(kTraceBlock),
		                   0, as_stream(s), tris, (uint32_t)n_tris, nodes, (uint32_t)(2 * n_tris - 1),
		                   ray_orig, ray_dir, max_t, (int)n, hits, stats, err);
	else
		hipLaunchKernelGGL(trace_cis_kernel<false>, dim3(grid_for(n, kTraceBlock)), dim3(kTraceBlock),
		                   0, as_stream(s), tris, (uint32_t)n_tris, nodes, (uint32_t)(2 * n_tris - 1),
		                   ray_orig, ray_dir, max_t, (int)n, hits, stats, err);
	RTA_LAUNCH_CHECK("trace_cis_kernel");
	return RTA_OK;
}

int rta_bvh_pack(rta_node_pair *pairs, float *tri4, uint32_t *flags, const rta_node *nodes,
                 int n_nodes, const rta_triangle *tris, int n_tris, rta_stream s) {
	if (!pairs || !tri4 || !flags || !nodes || !tris || n_nodes <= 0 || n_tris <= 0)
		return invalid("rta_bvh_pack: bad arguments");
	hipLaunchKernelGGL(pack_tris_kernel, dim3(grid_for(n_tris, 256)), dim3(256), 0, as_stream(s),
	                   (float4 *)tri4, tris, flags, n_tris);
	RTA_LAUNCH_CHECK("pack_tris_kernel");
	hipLaunchKernelGGL(pack_pairs_kernel, dim3(grid_for(n_nodes, 256)), dim3(256), 0, as_stream(s),
	                   pairs, (float4 *)tri4, flags, nodes, n_nodes);
	RTA_LAUNCH_CHECK("pack_pairs_kernel");
	return RTA_OK;
}

int rta_trace_cis_packed(const float *tri4, const rta_node_pair *pairs, const uint32_t *flags,
                         uint32_t root_link, uint32_t root_index, const float *ray_orig,
                         const float *ray_dir, const float *max_t, int n_rays,
                         const uint32_t *ray_count, rta_hit *hits, uint32_t *err, rta_stream s) {
	return trace_packed<false>("rta_trace_cis_packed", tri4, pairs, flags, root_link, root_index,
	                           ray_orig, ray_dir, max_t, n_rays, ray_count, hits, err,
	                           as_stream(s));
}

int rta_trace_shadow_cis_packed(const float *tri4, const rta_node_pair *pairs,
                                const uint32_t *flags, uint32_t root_link, uint32_t root_index,
                                const float *ray_orig, const float *ray_dir, const float *max_t,
                                int n_rays, const uint32_t *ray_count, rta_hit *hits,
                                uint32_t *err, rta_stream s) {
	return trace_packed<true>("rta_trace_shadow_cis_packed", tri4, pairs, flags, root_link,
	                          root_index, ray_orig, ray_dir, max_t, n_rays, ray_count, hits, err,
	                          as_stream(s));
}

int rta_trace_cis_persistent(const float *tri4, const rta_node_pair *pairs,
                             const uint32_t *flags, uint32_t root_link, uint32_t root_index,
                             const float *ray_orig, const float *ray_dir, const float *max_t,
                             int n_rays, const uint32_t *ray_count, rta_hit *hits, uint32_t *err,
                             uint32_t *queue, const rta_trace_config *cfg, rta_stream s) {
	RTA_FAULT_POINT("rta_trace_cis_persistent");
	return trace_persistent<false>("rta_trace_cis_persistent", tri4, pairs, flags, root_link,
	                               root_index, ray_orig, ray_dir, max_t, n_rays, ray_count, hits,
	                               err, queue, cfg, as_stream(s));
}

int rta_trace_cis_persistent_ovf(const float *tri4, const rta_node_pair *pairs,
                                 const uint32_t *flags, uint32_t root_link, uint32_t root_index,
                                 const float *ray_orig, const float *ray_dir, const float *max_t,
                                 int n_rays, const uint32_t *ray_count, rta_hit *hits,
                                 uint32_t *err, uint32_t *queue, const rta_trace_config *cfg,
                                 uint32_t *overflow, rta_stream s) {
	RTA_FAULT_POINT("rta_trace_cis_persistent_ovf");
	if (!overflow) return invalid("rta_trace_cis_persistent_ovf: NULL overflow list");
	return trace_persistent<false>("rta_trace_cis_persistent_ovf", tri4, pairs, flags, root_link,
	                               root_index, ray_orig, ray_dir, max_t, n_rays, ray_count, hits,
	                               err, queue, cfg, as_stream(s), overflow);
}

int rta_trace_shadow_cis_persistent_ovf(const float *tri4, const rta_node_pair *pairs,
                                        const uint32_t *flags, uint32_t root_link,
                                        uint32_t root_index, const float *ray_orig,
                                        const float *ray_dir, const float *max_t, int n_rays,
                                        const uint32_t *ray_count, rta_hit *hits, uint32_t *err,
                                        uint32_t *queue, const rta_trace_config *cfg,
                                        uint32_t *overflow, rta_stream s) {
	if (!overflow) return invalid("rta_trace_shadow_cis_persistent_ovf: NULL overflow list");
	return trace_persistent<true>("rta_trace_shadow_cis_persistent_ovf", tri4, pairs, flags,
	                              root_link, root_index, ray_orig, ray_dir, max_t, n_rays,
	                              ray_count, hits, err, queue, cfg, as_stream(s), overflow);
}

int rta_trace_cis_persistent_seeded(const float *tri4, const rta_node_pair *pairs,
                                    const uint32_t *flags, uint32_t root_link,
                                    uint32_t root_index, const float *ray_orig,
                                    const float *ray_dir, const float *max_t, int n_rays,
                                    const uint32_t *ray_count, const rta_treelet *table,
                                    const uint32_t *count, const int32_t *top, rta_hit *hits,
                                    uint32_t *err, uint32_t *queue, const rta_trace_config *cfg,
                                    rta_stream s) {
	const char *name = "rta_trace_cis_persistent_seeded";
	if (!tri4 || !pairs || !flags || !ray_orig || !ray_dir || !max_t || !hits || !err ||
	    !queue || !table || !count || !top || n_rays <= 0)
		return invalid("rta_trace_cis_persistent_seeded: bad arguments");
	rta_trace_config c;
	int rc = resolve_config(name, cfg, c);
	if (rc == RTA_OK) rc = require_full_stack(name, c);
	if (rc != RTA_OK) return rc;
	const uint32_t root_word = link_inner(root_link) ? 1u : (root_index << 1);
	SeedArgs seed;
	seed.table = table;
	seed.count = count;
	seed.top = top;
	const hipStream_t st = as_stream(s);
	switch (c.block) {
	case 64:
		return launch_persistent<64, false, false, true>(
		    tri4, pairs, flags, root_word, ray_orig, ray_dir, max_t, n_rays, ray_count, hits, err,
		    queue, c.refill_min, c.steps, c.blocks_per_cu, c.leaf_min, nullptr, 0, nullptr, st, seed);
	case 128:
		return launch_persistent<128, false, false, true>(
		    tri4, pairs, flags, root_word, ray_orig, ray_dir, max_t, n_rays, ray_count, hits, err,
		    queue, c.refill_min, c.steps, c.blocks_per_cu, c.leaf_min, nullptr, 0, nullptr, st, seed);
	default:
		return launch_persistent<256, false, false, true>(
		    tri4, pairs, flags, root_word, ray_orig, ray_dir, max_t, n_rays, ray_count, hits, err,
		    queue, c.refill_min, c.steps, c.blocks_per_cu, c.leaf_min, nullptr, 0, nullptr, st, seed);
	}
}

int rta_trace_shadow_cis_persistent(const float *tri4, const rta_node_pair *pairs,
                                    const uint32_t *flags, uint32_t root_link,
                                    uint32_t root_index, const float *ray_orig,
                                    const float *ray_dir, const float *max_t, int n_rays,
                                    const uint32_t *ray_count, rta_hit *hits, uint32_t *err,
                                    uint32_t *queue, const rta_trace_config *cfg, rta_stream s) {
	return trace_persistent<true>("rta_trace_shadow_cis_persistent", tri4, pairs, flags,
	                              root_link, root_index, ray_orig, ray_dir, max_t, n_rays,
	                              ray_count, hits, err, queue, cfg, as_stream(s));
}

int rta_treelet_cut(const rta_node_pair *pairs, uint32_t root_link, uint32_t root_index,
                    const uint32_t *ranges, int n_tris, int max_treelets, rta_treelet *table,
                    uint32_t *count, rta_stream s) {
	if (!pairs || !ranges || !table || !count || n_tris <= 0 || max_treelets < 1 ||
	    max_treelets > kMaxTreelets)
		return invalid("rta_treelet_cut: bad arguments (max_treelets in 1..1024)");
	const uint32_t root_word = link_inner(root_link) ? 1u : (root_index << 1);
	hipLaunchKernelGGL(treelet_cut_kernel, dim3(1), dim3(kMaxTreelets), 0, as_stream(s),
	                   (const float4 *)pairs, root_word, (const uint2 *)ranges, (uint32_t)n_tris,
	                   max_treelets, table, count);
	RTA_LAUNCH_CHECK("treelet_cut_kernel");
	return RTA_OK;
}

int rta_treelet_labels(const rta_hit *hits, int n, const rta_treelet *table,
                       const uint32_t *count, int32_t *label, float *target, rta_stream s) {
	if (!hits || !table || !count || !label || !target || n < 0)
		return invalid("rta_treelet_labels: bad arguments");
	if (n == 0) return RTA_OK;
	hipLaunchKernelGGL(treelet_labels_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(s),
	                   hits, n, table, count, label, target);
	RTA_LAUNCH_CHECK("treelet_labels_kernel");
	return RTA_OK;
}

int rta_trace_seed_treelets(const float *tri4, const rta_node_pair *pairs, const uint32_t *flags,
                            const rta_treelet *table, const uint32_t *count,
                            const float *ray_orig, const float *ray_dir, const float *max_t,
                            int n_rays, const int32_t *top, int candidates, float *t_seed,
                            uint32_t *err, rta_stream s) {
	if (!tri4 || !pairs || !flags || !table || !count || !ray_orig || !ray_dir || !max_t ||
	    !top || !t_seed || !err || n_rays <= 0 || candidates < 1 || candidates > kSeedTop)
		return invalid("rta_trace_seed_treelets: bad arguments (candidates in 1..4)");
	hipLaunchKernelGGL(seed_treelets_kernel, dim3(grid_for(n_rays, kTraceBlock)),
	                   dim3(kTraceBlock), 0, as_stream(s), (const float4 *)tri4,
	                   (const float4 *)pairs, flags, table, count, ray_orig, ray_dir, max_t, n_rays,
	                   top, candidates, t_seed, err);
	RTA_LAUNCH_CHECK("seed_treelets_kernel");
	return RTA_OK;
}

int rta_seed_max_t(float *out, const float *max_t, const float *t_seed, int n, rta_stream s) {
	if (!out || !max_t || !t_seed || n <= 0) return invalid("rta_seed_max_t: bad arguments");
	hipLaunchKernelGGL(seed_max_t_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(s), out,
	                   max_t, t_seed, n);
	RTA_LAUNCH_CHECK("seed_max_t_kernel");
	return RTA_OK;
}

int rta_batch_pack(rta_batch b, rta_node_pair *pairs, float *tri4, uint32_t *flags,
                   const rta_node *nodes, const rta_triangle *tris, rta_stream s) {
	if (!b || !pairs || !tri4 || !flags || !nodes || !tris)
		return invalid("rta_batch_pack: NULL argument");
	const BatchPlan &p = *(const BatchPlan *)b;
	hipStream_t st = as_stream(s);
	const int rc = rta_hip_fill_u32(flags, 0u, (size_t)p.n_meshes, s);
	if (rc != RTA_OK) return rc;
	hipLaunchKernelGGL(batch_pack_tris_kernel, dim3((unsigned)p.tri_blocks), dim3(256), 0, st,
	                   p.d_items + p.tri_blocks_off, p.d_meshes, (float4 *)tri4, tris);
	RTA_LAUNCH_CHECK("batch_pack_tris_kernel");
	hipLaunchKernelGGL(batch_pack_pairs_kernel, dim3((unsigned)p.node_blocks), dim3(256), 0, st,
	                   p.d_items + p.node_blocks_off, p.d_meshes, pairs, (float4 *)tri4, flags,
	                   nodes, p.d_root_words);
	RTA_LAUNCH_CHECK("batch_pack_pairs_kernel");
	return RTA_OK;
}

int rta_batch_trace_persistent(rta_batch b, const float *tri4, const rta_node_pair *pairs,
                               const uint32_t *flags, const float *ray_orig, const float *ray_dir,
                               const float *max_t, rta_hit *hits, uint32_t *err, uint32_t *queue,
                               int any_hit, const rta_trace_config *cfg, rta_stream s) {
	if (!b || !tri4 || !pairs || !flags || !ray_orig || !ray_dir || !max_t || !hits || !err ||
	    !queue)
		return invalid("rta_batch_trace_persistent: NULL argument");
	const BatchPlan &p = *(const BatchPlan *)b;
	if (p.total_rays <= 0) return invalid("rta_batch_trace_persistent: the batch has no rays");
	rta_trace_config c;
	const int rc = resolve_config("rta_batch_trace_persistent", cfg, c);
	if (rc != RTA_OK) return rc;
	hipStream_t st = as_stream(s);
	if (any_hit)
		return launch_persistent_cfg<true, true>(c, tri4, pairs, flags, 0u, ray_orig, ray_dir,
		                                         max_t, (int)p.total_rays, nullptr, hits, err,
		                                         queue, p.d_meshes, p.n_meshes, p.d_root_words, st,
		                                         p.d_ovf);
	return launch_persistent_cfg<false, true>(c, tri4, pairs, flags, 0u, ray_orig, ray_dir, max_t,
	                                          (int)p.total_rays, nullptr, hits, err, queue,
	                                          p.d_meshes, p.n_meshes, p.d_root_words, st, p.d_ovf);
}

// Coordinate descent over the launch shape: starting from `*best` (or the
// defaults), each round tries every candidate of one parameter at a time
// (leaf_min, refill_min, steps, block, stack) with the others fixed and keeps the
// fastest, timing each candidate as the best of `reps` launches.  Results do
// not depend on the shape, so the hits of the last launch are valid.
int rta_batch_autotune(rta_batch b, const float *tri4, const rta_node_pair *pairs,
                       const uint32_t *flags, const float *ray_orig, const float *ray_dir,
                       const float *max_t, rta_hit *hits, uint32_t *err, uint32_t *queue,
                       int any_hit, int rounds, int reps, rta_trace_config *best, float *best_ms,
                       rta_stream s) {
	if (!best) return invalid("rta_batch_autotune: best is NULL");
	if (rounds < 1 || reps < 1) return invalid("rta_batch_autotune: rounds and reps must be >= 1");
	rta_trace_config cur;
	int rc = resolve_config("rta_batch_autotune", best->block ? best : nullptr, cur);
	if (rc != RTA_OK) return rc;
	hipStream_t st = as_stream(s);
	hipEvent_t e0, e1;
	RTA_HIP_TRY(hipEventCreate(&e0));
	hipError_t he = hipEventCreate(&e1);
	if (he != hipSuccess) {
		(void)hipEventDestroy(e0);
		return hip_status(he, "hipEventCreate");
	}
	auto measure = [&](const rta_trace_config &c, float &ms) -> int {
		ms = 1e30f;
		for (int r = 0; r < reps; ++r) {
			hipError_t e = hipEventRecord(e0, st);
			if (e != hipSuccess) return hip_status(e, "hipEventRecord");
			int k = rta_batch_trace_persistent(b, tri4, pairs, flags, ray_orig, ray_dir, max_t, hits,
			                                   err, queue, any_hit, &c, s);
			if (k != RTA_OK) return k;
			e = hipEventRecord(e1, st);
			if (e == hipSuccess) e = hipEventSynchronize(e1);
			float t = 0;
			if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
			if (e != hipSuccess) return hip_status(e, "autotune timing");
			ms = std::min(ms, t);
		}
		return RTA_OK;
	};
	float cur_ms = 0;
	rc = measure(cur, cur_ms);
	static const int leafs[] = {2, 4, 6, 8, 12};
	static const int refills[] = {4, 8, 16, 24, 32};
	static const int stepss[] = {8, 16, 24, 32, 64};
	static const int blocks[] = {64, 128};
	static const int stacks[] = {16, 24, kStack};
	for (int round = 0; round < rounds && rc == RTA_OK; ++round) {
		for (int param = 0; param < 5 && rc == RTA_OK; ++param) {
			const int *vals = param == 0 ? leafs : param == 1 ? refills : param == 2 ? stepss
			                  : param == 3 ? blocks : stacks;
			const int nv = param == 3 ? 2 : (param == 4 ? 3 : 5);
			for (int v = 0; v < nv && rc == RTA_OK; ++v) {
				rta_trace_config c = cur;
				int &field = param == 0 ? c.leaf_min : param == 1 ? c.refill_min
				             : param == 2 ? c.steps : param == 3 ? c.block : c.stack;
				if (field == vals[v]) continue;
				field = vals[v];
				float ms = 0;
				rc = measure(c, ms);
				if (rc == RTA_OK && ms < cur_ms) { cur = c; cur_ms = ms; }
			}
		}
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (rc != RTA_OK) return rc;
	*best = cur;
	if (best_ms) *best_ms = cur_ms;
	return RTA_OK;
}

int rta_bounce_work_bytes(int n_rays, size_t *bytes) {
	if (n_rays <= 0 || !bytes) return invalid("rta_bounce_work_bytes: bad arguments");
	*bytes = (size_t)grid_for(n_rays, kBounceBlock) * 4;
	return RTA_OK;
}

}  // extern "C"

namespace rta_hip {

template <bool SHADOW>
int emit_rays(const char *name, float *out_orig, float *out_dir, float *out_maxt,
              uint32_t *src_index, uint32_t *count, const rta_hit *hits, const rta_triangle *tris,
              const float *ray_orig, const float *ray_dir, int n_rays, uint64_t seed,
              const float *light, void *work, size_t work_bytes, hipStream_t st) {
	if (!out_orig || !out_dir || !out_maxt || !src_index || !count || !hits || !tris ||
	    !ray_orig || !ray_dir || !work || n_rays <= 0 || (SHADOW && !light)) {
		set_error(std::string(name) + ": bad arguments");
		return RTA_E_INVALID;
	}
	size_t need = 0;
	rta_bounce_work_bytes(n_rays, &need);
	if (work_bytes < need) {
		set_error(std::string(name) + ": work buffer too small");
		return RTA_E_WORKSPACE;
	}
	const unsigned nb = grid_for(n_rays, kBounceBlock);
	uint32_t *bc = reinterpret_cast<uint32_t *>(work);
	hipLaunchKernelGGL(bounce_count_kernel, dim3(nb), dim3(kBounceBlock), 0, st, hits, n_rays, bc);
	RTA_LAUNCH_CHECK("bounce_count_kernel");
	hipLaunchKernelGGL(bounce_scan_kernel, dim3(1), dim3(1024), 0, st, bc, (int)nb, count);
	RTA_LAUNCH_CHECK("bounce_scan_kernel");
	const float lx = SHADOW ? light[0] : 0.0f, ly = SHADOW ? light[1] : 0.0f,
	            lz = SHADOW ? light[2] : 0.0f;
	hipLaunchKernelGGL(bounce_emit_kernel<SHADOW>, dim3(nb), dim3(kBounceBlock), 0, st, out_orig,
	                   out_dir, out_maxt, src_index, bc, hits, tris, ray_orig, ray_dir, n_rays,
	                   seed, lx, ly, lz);
	RTA_LAUNCH_CHECK("bounce_emit_kernel");
	return RTA_OK;
}

}  // namespace rta_hip

extern "C" {

int rta_diffuse_bounce(float *out_orig, float *out_dir, float *out_maxt, uint32_t *src_index,
                       uint32_t *count, const rta_hit *hits, const rta_triangle *tris,
                       const float *ray_orig, const float *ray_dir, int n_rays, uint64_t seed,
                       void *work, size_t work_bytes, rta_stream s) {
	return emit_rays<false>("rta_diffuse_bounce", out_orig, out_dir, out_maxt, src_index, count,
	                        hits, tris, ray_orig, ray_dir, n_rays, seed, nullptr, work, work_bytes,
	                        as_stream(s));
}

int rta_shadow_rays(float *out_orig, float *out_dir, float *out_maxt, uint32_t *src_index,
                    uint32_t *count, const rta_hit *hits, const rta_triangle *tris,
                    const float *ray_orig, const float *ray_dir, int n_rays, const float *light,
                    void *work, size_t work_bytes, rta_stream s) {
	return emit_rays<true>("rta_shadow_rays", out_orig, out_dir, out_maxt, src_index, count, hits,
	                       tris, ray_orig, ray_dir, n_rays, 0, light, work, work_bytes,
	                       as_stream(s));
}

}  // extern "C"
