// Per-mesh bodies of the LBVH build stages (bbvh-cuda/lbvh-kernels.cu), as
// device functions over one mesh's base pointers.  The single-mesh kernels
// (lbvh.hip) and the batched kernels (batch.hip) both call them, so every
// mesh of a batch is built by exactly the single-mesh code.
#pragma once

#include "common.hpp"

namespace rta_hip {

constexpr int kInitBlock = 256;
constexpr int kTriWords = sizeof(rta_triangle) / 8;   // 13 x 8 bytes
constexpr float kBoundsInit = 100000000000.0f;        // lbvh-kernels.cu:519
constexpr int kBoundsBlock = 1024;
constexpr int kChunk = 1024;
constexpr int kSlots = 128;
constexpr int kMergeFan = 16;
constexpr uint32_t kStateInner = 0xffffffffu; // leaf count of an inner node in the climb

// ---------------------------------------------------------- wave helpers

__device__ inline float wave_min(float v) {
	for (int off = kWave / 2; off > 0; off >>= 1) v = sel_min(v, __shfl_xor(v, off));
	return v;
}
__device__ inline float wave_max(float v) {
	for (int off = kWave / 2; off > 0; off >>= 1) v = sel_max(v, __shfl_xor(v, off));
	return v;
}

// ------------------------------------------------ stage 1: initTriangles

// One block of kInitBlock threads handles triangles [block*256, +256) of a
// mesh (26,624 contiguous bytes): a coalesced copy src -> scratch staged
// through LDS (16-byte words where aligned), then each thread forms its triangle's box
// (lbvh-kernels.cu:141-153) and centroid ((a+b)+c)/3.0f (lbvh-kernels.cu:
// 60-63) from LDS, and the block reduces its boxes into one partial box.
__device__ inline void init_triangles_block(rta_triangle *__restrict__ scratch,
                                            const rta_triangle *__restrict__ src,
                                            rta_aabb *__restrict__ boxes,
                                            rta_float3 *__restrict__ centers,
                                            rta_aabb *__restrict__ partial, int n, int block) {
	__shared__ __align__(16) uint2 words[kInitBlock * kTriWords];
	__shared__ float red[6][kInitBlock / kWave];

	const int tid = threadIdx.x;
	const long long base = (long long)block * kInitBlock;
	const int count = (int)((n - base) < kInitBlock ? (n - base) : kInitBlock);
	const int nwords = count * kTriWords;
	const uint2 *gsrc = reinterpret_cast<const uint2 *>(src + base);
	uint2 *gdst = reinterpret_cast<uint2 *>(scratch + base);
	// 16-byte words when both ends are 16-byte aligned (a mesh of a batch
	// may start at an odd triangle, 8 bytes off), then the odd last word
	const bool wide = ((reinterpret_cast<uintptr_t>(gsrc) | reinterpret_cast<uintptr_t>(gdst)) &
	                   15u) == 0u;
	int w0 = 0;
	if (wide) {
		const uint4 *s4 = reinterpret_cast<const uint4 *>(gsrc);
		uint4 *d4 = reinterpret_cast<uint4 *>(gdst);
		uint4 *l4 = reinterpret_cast<uint4 *>(words);
		const int n4 = nwords / 2;
		for (int w = tid; w < n4; w += kInitBlock) {
			uint4 v = s4[w];
			l4[w] = v;
			d4[w] = v;
		}
		w0 = 2 * n4;
	}
	for (int w = w0 + tid; w < nwords; w += kInitBlock) {
		uint2 v = gsrc[w];
		words[w] = v;
		gdst[w] = v;
	}
	__syncthreads();

	float mn[3] = {kBoundsInit, kBoundsInit, kBoundsInit};
	float mx[3] = {-kBoundsInit, -kBoundsInit, -kBoundsInit};
	if (tid < count) {
		const float *f = reinterpret_cast<const float *>(&words[tid * kTriWords]);
		float lo[3], hi[3], ce[3];
		for (int c = 0; c < 3; ++c) {
			float va = f[c], vb = f[3 + c], vc = f[6 + c];
			lo[c] = sel_min(sel_min(va, vb), vc);
			hi[c] = sel_max(sel_max(va, vb), vc);
			ce[c] = ((va + vb) + vc) / 3.0f;
			// a NaN box coordinate does not enter the scene box: with a NaN
			// operand the reference's MIN/MAX (reduce_max,
			// lbvh-kernels.cu:496-516) is neither associative nor
			// commutative, so its thrust::reduce result depends on the
			// reduction order; skipping NaNs makes the bounds the min/max
			// over the other coordinates in every order
			if (lo[c] == lo[c]) mn[c] = lo[c];
			if (hi[c] == hi[c]) mx[c] = hi[c];
		}
		rta_aabb b;
		b.min = {lo[0], lo[1], lo[2]};
		b.max = {hi[0], hi[1], hi[2]};
		boxes[base + tid] = b;
		centers[base + tid] = {ce[0], ce[1], ce[2]};
	}
	// block reduction of the boxes (min/max are exact; the order only
	// matters for the sign of a zero, which cannot change a Morton code)
	const int lane = tid % kWave, wv = tid / kWave;
	for (int c = 0; c < 3; ++c) {
		float a = wave_min(mn[c]);
		float b = wave_max(mx[c]);
		if (lane == 0) { red[c][wv] = a; red[3 + c][wv] = b; }
	}
	__syncthreads();
	if (tid == 0) {
		float r[6];
		for (int c = 0; c < 3; ++c) { r[c] = kBoundsInit; r[3 + c] = -kBoundsInit; }
		for (int w = 0; w < kInitBlock / kWave; ++w)
			for (int c = 0; c < 3; ++c) {
				r[c] = sel_min(r[c], red[c][w]);
				r[3 + c] = sel_max(r[3 + c], red[3 + c][w]);
			}
		rta_aabb p;
		p.min = {r[0], r[1], r[2]};
		p.max = {r[3], r[4], r[5]};
		*partial = p;
	}
}

// ----------------------------------------------- stage 2: voxel bounds

// One block of kBoundsBlock threads reduces a mesh's np partial boxes
// (calculateVoxelBounds, lbvh-kernels.cu:518-535).
__device__ inline void voxel_bounds_block(rta_aabb *__restrict__ bounds,
                                          const rta_aabb *__restrict__ partials, int np) {
	__shared__ float red[6][kBoundsBlock / kWave];
	float mn[3] = {kBoundsInit, kBoundsInit, kBoundsInit};
	float mx[3] = {-kBoundsInit, -kBoundsInit, -kBoundsInit};
	for (int i = threadIdx.x; i < np; i += kBoundsBlock) {
		rta_aabb p = partials[i];
		mn[0] = sel_min(mn[0], p.min.x); mn[1] = sel_min(mn[1], p.min.y); mn[2] = sel_min(mn[2], p.min.z);
		mx[0] = sel_max(mx[0], p.max.x); mx[1] = sel_max(mx[1], p.max.y); mx[2] = sel_max(mx[2], p.max.z);
	}
	const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
	for (int c = 0; c < 3; ++c) {
		float a = wave_min(mn[c]);
		float b = wave_max(mx[c]);
		if (lane == 0) { red[c][wv] = a; red[3 + c][wv] = b; }
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float r[6];
		for (int c = 0; c < 3; ++c) { r[c] = kBoundsInit; r[3 + c] = -kBoundsInit; }
		for (int w = 0; w < kBoundsBlock / kWave; ++w)
			for (int c = 0; c < 3; ++c) {
				r[c] = sel_min(r[c], red[c][w]);
				r[3 + c] = sel_max(r[3 + c], red[3 + c][w]);
			}
		rta_aabb b;
		b.min = {r[0], r[1], r[2]};
		b.max = {r[3], r[4], r[5]};
		*bounds = b;
	}
}

// ------------------------------------------------ stage 3: Morton codes

// The code of triangle i of a mesh.  d = GRIDSIZE/(max-min) is formed per
// thread with the same correctly rounded division the reference does once on
// the host (lbvh-kernels.cu:414-417), so every thread sees the identical d.
__device__ inline uint32_t morton_of(const rta_float3 *__restrict__ centers,
                                     const rta_aabb *__restrict__ bounds, int i) {
	const rta_aabb vb = *bounds;
	const float dx = 1024.0f / (vb.max.x - vb.min.x);
	const float dy = 1024.0f / (vb.max.y - vb.min.y);
	const float dz = 1024.0f / (vb.max.z - vb.min.z);
	const rta_float3 c = centers[i];
	const uint32_t x = f2u_rz_sat((c.x - vb.min.x) * dx);
	const uint32_t y = f2u_rz_sat((c.y - vb.min.y) * dy);
	const uint32_t z = f2u_rz_sat((c.z - vb.min.z) * dz);
	return morton_encode_10bit(x, y, z);
}

// ----------------------------------------- stage 5: Karras hierarchy

// The sorted codes of one mesh as the hierarchy searches read them:
// positions [lo, hi) from a copy in LDS, the rest from global memory.  Most
// probes of a node's searches fall within a few times its range length of
// the node, so a block that stages its own positions plus a halo on each
// side answers nearly all of them from LDS.
constexpr int kKarrasBlock = 256;
constexpr int kKarrasHalo = 256;   // 1024 measured slower: staging outweighs the probes saved
constexpr int kKarrasWindow = kKarrasBlock + 2 * kKarrasHalo;
struct CodeWindow {
	const uint32_t *__restrict__ codes;
	const uint32_t *lds;
	int lo, hi;
	__device__ uint32_t operator()(int j) const {
		return (j >= lo && j < hi) ? lds[j - lo] : codes[j];
	}
};

// Stages positions [base - halo, base + block + halo) of [0, n) in lds
// (kKarrasWindow words) when the block holds inner nodes (base < n - 1);
// every thread of the block must call it.
__device__ inline CodeWindow stage_codes(uint32_t *lds, const uint32_t *__restrict__ codes, int n,
                                         int base) {
	CodeWindow w{codes, lds, 0, 0};
	if (base >= n - 1) return w;
	w.lo = base - kKarrasHalo > 0 ? base - kKarrasHalo : 0;
	w.hi = base + kKarrasBlock + kKarrasHalo < n ? base + kKarrasBlock + kKarrasHalo : n;
	for (int j = w.lo + (int)threadIdx.x; j < w.hi; j += kKarrasBlock) lds[j - w.lo] = codes[j];
	__syncthreads();
	return w;
}

// Common prefix length of the augmented keys (code, position):
// clz(ci^cj), or 32+clz(i^j) for equal codes, -1 outside [0,n)
// (__cpl, lbvh-kernels.cu:84-92); ci is the code of position i.
__device__ inline int cpl(int i, uint32_t ci, int j, const CodeWindow &w, int n) {
	if (j < 0 || j >= n) return -1;
	const uint32_t cj = w(j);
	if (ci == cj) return __clz((int)((uint32_t)i ^ (uint32_t)j)) + 32;
	return __clz((int)(ci ^ cj));
}

// Node i of a mesh of n triangles (k::build_tree_karras, lbvh-kernels.cu:
// 186-256).  The direction/range/split searches are the reference's,
// including its ceil-halving split search that always runs in +1 direction
// from the lower end of the range.  Inner nodes also clear their visit
// counter and record their sorted range [from, to] for the bottom-up pass.
__device__ inline void karras_node(rta_node *__restrict__ nodes, uint32_t *__restrict__ parents,
                                   uint32_t *__restrict__ visit, uint2 *__restrict__ ranges,
                                   const CodeWindow &w, int n, int i) {
	uint2 *links = reinterpret_cast<uint2 *>(&nodes[i]);
	if (i >= n - 1) {
		// leaf: tris = i-(n-1), elems = 1, leaf bit clear
		*links = make_uint2(1u << 1, (uint32_t)(i - (n - 1)));
		return;
	}
	const uint32_t ci = w(i);
	int d = (cpl(i, ci, i + 1, w, n) - cpl(i, ci, i - 1, w, n)) >= 0 ? 1 : -1;
	const int cpl_min = cpl(i, ci, i - d, w, n);
	int l_max = 2;
	while (cpl(i, ci, i + l_max * d, w, n) > cpl_min) l_max *= 2;
	int l = 0;
	for (int t = l_max / 2; t >= 1; t /= 2)
		if (cpl(i, ci, i + (l + t) * d, w, n) > cpl_min) l += t;
	const int j = i + l * d;
	const int from = j < i ? j : i;
	const int to = j < i ? i : j;

	const uint32_t cfrom = w(from);
	const int cpl_node = cpl(from, cfrom, to, w, n);
	int s = 0;
	for (int t = (l / 2 + l % 2); t >= 1; t = (t / 2 + t % 2)) {
		if (cpl(from, cfrom, from + (s + t), w, n) > cpl_node) s += t;
		if (t <= 1) break;
	}
	const int mid = from + s;
	int left = mid, right = mid + 1;
	if (from == mid) left += n - 1;
	if (to == mid + 1) right += n - 1;
	*links = make_uint2(((uint32_t)left << 1) | 1u, (uint32_t)right);
	parents[left] = (uint32_t)i;
	parents[right] = (uint32_t)i;
	visit[i] = 0u;
	ranges[i] = make_uint2((uint32_t)from, (uint32_t)to);
}

// ----------------------------------------- stage 6: bottom-up boxes

// Bottom-up boxes (k::build_bounding_boxes_karras2, lbvh-kernels.cu:258-341),
// with the hierarchy of k::build_tree_karras (:186-256) emitted by the same
// climb.  The hierarchy is the binary radix tree of the augmented keys
// (code, position), which is unique, so it can be grown from the leaves: a
// completed node covering sorted positions [l, r] has its parent on the
// side of the longer common prefix with its neighbour -- split r (the node
// is the left child) when delta(r) > delta(l-1), else split l-1 (the right
// child), with delta(i) the common prefix length of positions i and i+1
// (cpl above) and delta(-1) = delta(n-1) = -1; two neighbours never tie.
// Karras numbering follows: the children of split s are nodes s and s+1
// (leaves when their range is one position), so a node's index is r as a
// left child, l as a right child, and 0 for the root.
//
// At each split the first arriving child stops and the second forms the
// parent: the union MIN(left, right) / MAX(left, right) with the
// reference's `(a<b)?a:b`, the whole 32-byte node, its range and (packed
// build) its pair record, and climbs on.  Merge rule: both children are
// leaves (original or merged) and their representative codes are equal
// (codes[s] == codes[s+1], i.e. delta(s) >= 32; `code == old` in the
// reference) -> the parent becomes one leaf with tris = min, elems = sum
// (lbvh-kernels.cu:302-330).  Deterministic form of the reference: a visit
// counter per split decides first/second arrival (the reference sums
// representative codes into the flag, so a first arrival with code 0 also
// stops the second thread, :288-291), and the union order is fixed (the
// reference leaves the sign of a zero to arrival order).
//
// Hierarchical schedule.  A subtree covers a contiguous range of sorted
// positions, so all communication inside a chunk of consecutive leaves
// stays in one workgroup:
//  level 1: a 1024-thread block owns 1024 leaves; the deltas, the visit
//           counters and the waiting children live in LDS; a climb stops
//           when its parent's split lies outside the chunk, and the
//           completed nodes whose parents stay incomplete form the chunk's
//           frontier (at most 2 per tree level, so <= 126 for depth <= 63);
//  level k: one block merges the frontiers of kMergeFan level-(k-1) chunks,
//           its visit counters and waiting children in LDS again (a table
//           keyed by split), until one block covers all leaves and reaches
//           the root.
// No step needs device-scope cache maintenance; results do not depend on
// the schedule.

__device__ inline rta_aabb union_lr(const rta_aabb &lb, const rta_aabb &rb) {
	rta_aabb u;
	u.min.x = sel_min(lb.min.x, rb.min.x);
	u.min.y = sel_min(lb.min.y, rb.min.y);
	u.min.z = sel_min(lb.min.z, rb.min.z);
	u.max.x = sel_max(lb.max.x, rb.max.x);
	u.max.y = sel_max(lb.max.y, rb.max.y);
	u.max.z = sel_max(lb.max.z, rb.max.z);
	return u;
}

// The traversal layout written by the build itself (rta_lbvh_build_packed):
// per inner node x a 64-byte child-pair record pairs[x] (both children's
// boxes SoA, their words -- (x << 1) | 1 for an inner child, first
// triangle << 1 for a leaf -- and their node ids), per triangle three
// float4 of tri4 (a, a - b, a - c; the w of the first is nonzero on the last
// triangle of a leaf), and flags bit 0 when some box coordinate leaves the
// exact-division range.  The same bytes as rta_bvh_pack over the finished
// nodes.  All pointers null: the plain build.
struct PackOut {
	rta_node_pair *pairs = nullptr;
	float4 *tri4 = nullptr;
	uint32_t *flags = nullptr;
};

__device__ inline bool box_fast(const rta_aabb &b) {
	return coord_fast_ok(b.min.x) && coord_fast_ok(b.min.y) && coord_fast_ok(b.min.z) &&
	       coord_fast_ok(b.max.x) && coord_fast_ok(b.max.y) && coord_fast_ok(b.max.z);
}

// pairs[x] from its children's boxes, words and ids; returns false when a
// box coordinate leaves the exact-division range
__device__ inline bool write_pair(rta_node_pair *pairs, uint32_t x, const rta_aabb &lb,
                                  const rta_aabb &rb, uint32_t lword, uint32_t rword, uint32_t lid,
                                  uint32_t rid) {
	rta_node_pair p;
	p.lo_x[0] = lb.min.x; p.lo_x[1] = rb.min.x;
	p.lo_y[0] = lb.min.y; p.lo_y[1] = rb.min.y;
	p.lo_z[0] = lb.min.z; p.lo_z[1] = rb.min.z;
	p.hi_x[0] = lb.max.x; p.hi_x[1] = rb.max.x;
	p.hi_y[0] = lb.max.y; p.hi_y[1] = rb.max.y;
	p.hi_z[0] = lb.max.z; p.hi_z[1] = rb.max.z;
	p.link[0] = lword; p.link[1] = rword;
	p.index[0] = lid; p.index[1] = rid;
	pairs[x] = p;
	return box_fast(lb) && box_fast(rb);
}

// a merge joins two leaves: the left one's last triangle no longer ends a
// leaf (the permute wrote every triangle as a leaf end)
__device__ inline void clear_leaf_end(float4 *tri4, uint32_t pos) {
	tri4[3 * (size_t)pos].w = 0.0f;
}

// delta(i): common prefix length of the augmented keys at sorted positions
// i and i+1, -1 outside [0, n-1)
__device__ inline int delta_at(const uint32_t *__restrict__ codes, long long i, int n) {
	if (i < 0 || i >= (long long)n - 1) return -1;
	const uint32_t a = codes[i], b = codes[i + 1];
	if (a == b) return __clz((int)((uint32_t)i ^ (uint32_t)(i + 1))) + 32;
	return __clz((int)(a ^ b));
}

// A completed node during the climb.
struct ClimbNode {
	uint32_t l, r;     // sorted range
	int dl, dr;        // delta(l - 1), delta(r)
	rta_aabb box;
	uint32_t elems;    // kStateInner for an inner node, else the leaf's triangle count
	uint32_t tris;     // the leaf's first triangle
};

// its parent's split, and its own index there (r as a left child, l as a
// right child); the root (dl = dr = -1) has neither
__device__ inline bool is_left_child(const ClimbNode &c) { return c.dr > c.dl; }
__device__ inline uint32_t node_index(const ClimbNode &c, uint32_t leaf0) {
	if (c.l == c.r) return leaf0 + c.l;
	return is_left_child(c) ? c.r : c.l;
}

// a frontier entry: index, range, deltas (+1, a byte each)
__device__ inline uint4 frontier_entry(const ClimbNode &c, uint32_t leaf0) {
	return make_uint4(node_index(c, leaf0), c.l, c.r,
	                  (uint32_t)(c.dl + 1) | ((uint32_t)(c.dr + 1) << 8));
}

// the whole node: link words and box, as two 16-byte stores
__device__ inline void store_node(rta_node *node, uint2 links, const rta_aabb &b) {
	uint4 *o = reinterpret_cast<uint4 *>(node);
	o[0] = make_uint4(links.x, links.y, __float_as_uint(b.min.x), __float_as_uint(b.min.y));
	o[1] = make_uint4(__float_as_uint(b.min.z), __float_as_uint(b.max.x), __float_as_uint(b.max.y),
	                  __float_as_uint(b.max.z));
}

// leaf state of a finished node from its link words
__device__ inline void node_state(uint2 links, uint32_t &elems, uint32_t &tris) {
	elems = link_inner(links.x) ? kStateInner : links.x >> 1;
	tris = links.y;
}

// The parent of lc = [l, s] and rc = [s+1, r]: its whole node, range and,
// after a merge, the cleared leaf-end mark; returns it, its index in x.
__device__ inline ClimbNode form_parent(const ClimbNode &lc, const ClimbNode &rc, uint32_t s,
                                        uint32_t leaf0, rta_node *__restrict__ nodes,
                                        uint2 *__restrict__ ranges, const PackOut &pk,
                                        uint32_t &x) {
	ClimbNode p;
	p.l = lc.l;
	p.r = rc.r;
	p.dl = lc.dl;
	p.dr = rc.dr;
	p.box = union_lr(lc.box, rc.box);
	const uint32_t lid = lc.l == lc.r ? leaf0 + s : s;
	const uint32_t rid = rc.l == rc.r ? leaf0 + s + 1u : s + 1u;
	const bool merge = lc.dr >= 32 && lc.elems != kStateInner && rc.elems != kStateInner;
	p.elems = merge ? lc.elems + rc.elems : kStateInner;
	p.tris = merge ? (lc.tris < rc.tris ? lc.tris : rc.tris) : 0u;
	x = node_index(p, leaf0);
	store_node(&nodes[x], merge ? make_uint2(p.elems << 1, p.tris) : make_uint2((lid << 1) | 1u, rid),
	           p.box);
	ranges[x] = make_uint2(p.l, p.r);
	if (merge && pk.tri4) clear_leaf_end(pk.tri4, s);
	return p;
}

// Level 1 for leaves [chunk*kChunk, +kChunk) of a mesh: every leaf's node,
// every node completed inside the chunk, and the frontier in slots[0..kSlots)
// with its size in *count.  The climbs touch LDS only: every child arriving
// at a split of the chunk writes its record (box, far end, the delta there,
// leaf state) under that split and its side, then counts its arrival; the
// second arrival reads its sibling's record and climbs on with the parent.
// Afterwards each split whose two children arrived writes the parent's
// node, range and pair record from the two records, one split per thread,
// and a split with one arrival gives a frontier entry.
constexpr uint16_t kInner16 = 0xffffu;      // record: an inner node
constexpr uint32_t kArriveRight = 0x10000u; // visit count of a right child (a left one adds 1)
__device__ inline void bottom_up_chunk(rta_node *__restrict__ nodes,
                                       const rta_aabb *__restrict__ boxes,
                                       const uint32_t *__restrict__ codes,
                                       const uint32_t *__restrict__ indices,
                                       uint2 *__restrict__ ranges,
                                       uint4 *__restrict__ slots, uint32_t *__restrict__ count,
                                       uint32_t *__restrict__ status, int n, int chunk,
                                       PackOut pk = PackOut{}) {
	// records per split base + j and side (0: the left child, 1: the right)
	__shared__ float rbox[2][6][kChunk];
	__shared__ uint16_t rfar[2][kChunk];     // the far end of the child's range, - base
	__shared__ uint16_t relems[2][kChunk];   // a leaf's triangle count, or kInner16
	__shared__ uint16_t rtris[2][kChunk];    // a leaf's first position, - base
	__shared__ int8_t rdfar[2][kChunk];      // delta at the far end
	__shared__ uint32_t lvisit[kChunk];      // arrivals per split
	__shared__ int8_t ldelta[kChunk + 1];    // delta(base + j - 1)
	__shared__ uint4 front[kSlots];
	__shared__ uint32_t nfront;

	const int t = threadIdx.x;
	const uint32_t base = (uint32_t)chunk * kChunk;
	const uint32_t leaf0 = (uint32_t)(n - 1);
	const uint32_t end = min(base + (uint32_t)kChunk, (uint32_t)n);
	const uint32_t cnt = end - base;
	const uint32_t k = base + t;
	lvisit[t] = 0;
	if (t == 0) nfront = 0;
	if ((uint32_t)t <= cnt) ldelta[t] = (int8_t)delta_at(codes, (long long)base + t - 1, n);
	if (t == 0 && cnt == (uint32_t)kChunk) ldelta[kChunk] = (int8_t)delta_at(codes, (long long)end - 1, n);
	const bool live = k < (uint32_t)n;
	ClimbNode c;
	if (live) {
		c.l = c.r = k;
		c.box = boxes[indices[k]];
		c.elems = 1u;
		c.tris = k;
		store_node(&nodes[leaf0 + k], make_uint2(1u << 1, k), c.box);
	}
	auto push = [&](const uint4 &e) {
		const uint32_t slot = atomicAdd(&nfront, 1u);
		if (slot < (uint32_t)kSlots) front[slot] = e;
		else atomicOr(status, (uint32_t)RTA_LBVH_FRONTIER_OVERFLOW);
	};
	// the child that arrived at split base + j on `side`
	auto record = [&](uint32_t j, int side) -> ClimbNode {
		ClimbNode a;
		a.box.min = {rbox[side][0][j], rbox[side][1][j], rbox[side][2][j]};
		a.box.max = {rbox[side][3][j], rbox[side][4][j], rbox[side][5][j]};
		const uint32_t far = base + rfar[side][j];
		const int dfar = rdfar[side][j], ds = ldelta[j + 1];
		if (side == 0) { a.l = far; a.r = base + j; a.dl = dfar; a.dr = ds; }
		else { a.l = base + j + 1u; a.r = far; a.dl = ds; a.dr = dfar; }
		const uint16_t e = relems[side][j];
		a.elems = e == kInner16 ? kStateInner : (uint32_t)e;
		a.tris = base + rtris[side][j];
		return a;
	};
	__syncthreads();

	if (live) {
		c.dl = ldelta[t];
		c.dr = ldelta[t + 1];
		for (int level = 0; level < 128; ++level) {
			if (c.dl < 0 && c.dr < 0) break;   // the root of the mesh
			const bool left = is_left_child(c);
			const uint32_t s = left ? c.r : c.l - 1u;
			if (s < base || s + 1u >= end) {
				push(frontier_entry(c, leaf0));   // the sibling starts in another chunk
				break;
			}
			const uint32_t j = s - base;
			const int side = left ? 0 : 1;
			rbox[side][0][j] = c.box.min.x; rbox[side][1][j] = c.box.min.y; rbox[side][2][j] = c.box.min.z;
			rbox[side][3][j] = c.box.max.x; rbox[side][4][j] = c.box.max.y; rbox[side][5][j] = c.box.max.z;
			rfar[side][j] = (uint16_t)((left ? c.l : c.r) - base);
			rdfar[side][j] = (int8_t)(left ? c.dl : c.dr);
			const bool inner = c.elems == kStateInner;
			relems[side][j] = inner ? kInner16 : (uint16_t)c.elems;
			rtris[side][j] = inner ? (uint16_t)0 : (uint16_t)(c.tris - base);
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
			const uint32_t old = atomicAdd(&lvisit[j], left ? 1u : kArriveRight);
			if (old == 0u) break;
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
			const ClimbNode sib = record(j, left ? 1 : 0);
			const ClimbNode &lc = left ? c : sib;
			const ClimbNode &rc = left ? sib : c;
			ClimbNode p;
			p.l = lc.l; p.r = rc.r; p.dl = lc.dl; p.dr = rc.dr;
			p.box = union_lr(lc.box, rc.box);
			const bool merge = lc.dr >= 32 && lc.elems != kStateInner && rc.elems != kStateInner;
			p.elems = merge ? lc.elems + rc.elems : kStateInner;
			p.tris = merge ? (lc.tris < rc.tris ? lc.tris : rc.tris) : 0u;
			c = p;
		}
	}
	__syncthreads();

	// write-out, per split s = base + t of the chunk: the parent when both
	// children arrived, a frontier entry when one did (its sibling reaches
	// beyond the chunk)
	bool slow = false;
	if ((uint32_t)t + 1u < cnt) {
		const uint32_t v = lvisit[t];
		if (v == 1u + kArriveRight) {
			const ClimbNode lc = record(t, 0), rc = record(t, 1);
			const uint32_t s = base + t;
			uint32_t x;
			form_parent(lc, rc, s, leaf0, nodes, ranges, pk, x);
			const bool merged = lc.dr >= 32 && lc.elems != kStateInner && rc.elems != kStateInner;
			if (!merged && pk.pairs) {
				const uint32_t lid = lc.l == lc.r ? leaf0 + s : s;
				const uint32_t rid = rc.l == rc.r ? leaf0 + s + 1u : s + 1u;
				const uint32_t lw = lc.elems == kStateInner ? (s << 1) | 1u : lc.tris << 1;
				const uint32_t rw = rc.elems == kStateInner ? ((s + 1u) << 1) | 1u : rc.tris << 1;
				slow = !write_pair(pk.pairs, x, lc.box, rc.box, lw, rw, lid, rid);
			}
		} else if (v != 0u) {
			push(frontier_entry(record(t, v == 1u ? 0 : 1), leaf0));
		}
	}
	if (pk.flags && __syncthreads_or(slow) && t == 0) atomicOr(pk.flags, 1u);
	__syncthreads();
	const uint32_t nf = min(nfront, (uint32_t)kSlots);
	if ((uint32_t)t < nf) slots[t] = front[t];
	if (t == 0) *count = nf;
}

// Level k >= 2 for one mesh: merges the frontier slots of the ng chunks of
// the previous level starting at in_slots/in_counts (kSlots entries per
// chunk, at most kMergeEntries in all), whose leaves span [lo, hi); the new
// frontier goes to out_slots / *out_count when given.  Thread t climbs from
// entries t, t + kMergeThreads, ... in turn (a climb never waits, so one
// thread may hold both children of a split).  The climbs touch LDS only:
// every arrival at a split writes its record (far end, the deltas there and
// at the split, box, leaf state) into its entry's record slot and then
// claims the split's entry of an open-addressing table with one 64-bit
// compare-and-swap of (split, record + 1); a climb that finds the split
// already there is the second and forms the parent from the first's
// record.  Only the first arrival stops, so a record is read only after its
// last write.  A block sees at most kMergeEntries distinct splits, so the
// table (half full at most) never runs out.  A child whose sibling lies
// beyond [lo, hi) -- its parent's split is outside, or its entry holds a
// single arrival at the end -- joins the new frontier.
constexpr int kMergeThreads = 1024;
constexpr int kMergeEntries = kMergeFan * kSlots;
constexpr int kTableBits = 12;
constexpr int kTable = 1 << kTableBits;              // >= 2 kMergeEntries
constexpr uint64_t kEmpty = ~0ull;                   // free table entry
// record word once both children arrived: kFormed, or kFormedInner | x for
// an inner parent x that needs its pair record
constexpr uint32_t kFormed = 0xffffffffu;
constexpr uint32_t kFormedInner = 0x80000000u;
static_assert(kTable >= 2 * kMergeEntries, "merge table at most half full");
__device__ inline void bottom_up_merge(rta_node *__restrict__ nodes, uint2 *__restrict__ ranges,
                                       const uint4 *__restrict__ in_slots,
                                       const uint32_t *__restrict__ in_counts, int ng,
                                       uint4 *__restrict__ out_slots,
                                       uint32_t *__restrict__ out_count,
                                       uint32_t *__restrict__ status, uint32_t lo, uint32_t hi,
                                       int n, PackOut pk = PackOut{}) {
	__shared__ uint32_t gcount[kMergeFan];
	// kEmpty, or (split << 32) | (the first arrival's record + 1, or the
	// formed mark)
	__shared__ uint64_t table[kTable];
	// records, one per entry
	__shared__ float rbox[6][kMergeEntries];
	__shared__ uint32_t rfar[kMergeEntries];     // far end of the child's range
	__shared__ uint32_t relems[kMergeEntries];
	__shared__ uint32_t rtris[kMergeEntries];
	__shared__ int8_t rdfar[kMergeEntries];      // delta at the far end
	__shared__ int8_t rdnear[kMergeEntries];     // delta at the split
	__shared__ uint8_t rleft[kMergeEntries];     // the child is the left one
	__shared__ uint4 front[kSlots];
	__shared__ uint32_t nfront;
	const int t = threadIdx.x;
	for (int i = t; i < kTable; i += blockDim.x) table[i] = kEmpty;
	// the groups' entry counts, one load per thread
	if (t < ng) gcount[t] = min(in_counts[t], (uint32_t)kSlots);
	if (t == 0) nfront = 0;
	__syncthreads();
	const uint32_t leaf0 = (uint32_t)(n - 1);
	auto push = [&](const uint4 &e) {
		const uint32_t slot = atomicAdd(&nfront, 1u);
		if (out_slots && slot < (uint32_t)kSlots) front[slot] = e;
		else atomicOr(status, (uint32_t)RTA_LBVH_FRONTIER_OVERFLOW);
	};
	uint32_t total = 0;
	for (int k = 0; k < ng; ++k) total += gcount[k];
	for (uint32_t ei = (uint32_t)t; ei < total; ei += kMergeThreads) {
		// entry ei: group g, position ei - before
		int g = 0;
		uint32_t before = 0;
		while (ei >= before + gcount[g]) before += gcount[g++];
		const uint4 e = in_slots[(size_t)g * kSlots + (ei - before)];
		ClimbNode c;
		c.l = e.y;
		c.r = e.z;
		c.dl = (int)(e.w & 0xffu) - 1;
		c.dr = (int)((e.w >> 8) & 0xffu) - 1;
		{
			const rta_node nd = nodes[e.x];
			c.box = nd.box;
			node_state(make_uint2(nd.type_left_elems, nd.right_tris), c.elems, c.tris);
		}
		for (int level = 0; level < 128; ++level) {
			if (c.dl < 0 && c.dr < 0) break;   // the root of the mesh
			const bool left = is_left_child(c);
			const uint32_t s = left ? c.r : c.l - 1u;
			if (s < lo || s + 1u >= hi) {
				push(frontier_entry(c, leaf0));
				break;
			}
			const uint32_t r = ei;
			rbox[0][r] = c.box.min.x; rbox[1][r] = c.box.min.y; rbox[2][r] = c.box.min.z;
			rbox[3][r] = c.box.max.x; rbox[4][r] = c.box.max.y; rbox[5][r] = c.box.max.z;
			rfar[r] = left ? c.l : c.r;
			rdfar[r] = (int8_t)(left ? c.dl : c.dr);
			rdnear[r] = (int8_t)(left ? c.dr : c.dl);
			relems[r] = c.elems;
			rtris[r] = c.tris;
			rleft[r] = left ? 1u : 0u;
			// LDS only (the node stores need not wait): the record before
			// the claim, the sibling's record after
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
			const uint64_t mine = ((uint64_t)s << 32) | (r + 1u);
			uint32_t h = (s * 2654435761u) >> (32 - kTableBits);
			uint64_t prev;
			while (true) {
				prev = atomicCAS(&table[h], kEmpty, mine);
				if (prev == kEmpty || (uint32_t)(prev >> 32) == s) break;
				h = (h + 1u) & (kTable - 1);
			}
			if (prev == kEmpty) break;   // the sibling forms the parent
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
			const uint32_t q = (uint32_t)prev - 1u;
			const int ds = left ? c.dr : c.dl;   // delta(s)
			ClimbNode sib;
			sib.box.min = {rbox[0][q], rbox[1][q], rbox[2][q]};
			sib.box.max = {rbox[3][q], rbox[4][q], rbox[5][q]};
			sib.elems = relems[q];
			sib.tris = rtris[q];
			if (left) { sib.l = s + 1u; sib.r = rfar[q]; sib.dl = ds; sib.dr = rdfar[q]; }
			else { sib.l = rfar[q]; sib.r = s; sib.dl = rdfar[q]; sib.dr = ds; }
			if (sib.l < lo || sib.r >= hi || sib.l > sib.r) {
				// cannot happen for sorted codes; guards the writes below
				atomicOr(status, (uint32_t)RTA_LBVH_FRONTIER_OVERFLOW);
				table[h] = ((uint64_t)s << 32) | kFormed;
				break;
			}
			uint32_t x;
			c = left ? form_parent(c, sib, s, leaf0, nodes, ranges, pk, x)
			         : form_parent(sib, c, s, leaf0, nodes, ranges, pk, x);
			table[h] = ((uint64_t)s << 32) |
			           (pk.pairs && c.elems == kStateInner ? kFormedInner | x : kFormed);
		}
	}
	__syncthreads();
	// Per table entry: a single arrival at a split (the sibling lies beyond
	// [lo, hi)) joins the frontier; an inner node formed here gets its pair
	// record from the finished nodes (written by this block's threads before
	// the barrier; inside the climb their stores held it up).
	// The thread's kPer entries are taken together, so that the node loads
	// of its pair records (links, then both children) are in flight at once.
	constexpr int kPer = kTable / kMergeThreads;
	static_assert(kTable % kMergeThreads == 0, "whole table entries per thread");
	uint32_t xs[kPer];
	bool inner[kPer];
	#pragma unroll
	for (int k = 0; k < kPer; ++k) {
		const uint64_t e = table[t + k * kMergeThreads];
		const uint32_t word = (uint32_t)e;
		inner[k] = e != kEmpty && word != kFormed && (word & kFormedInner);
		xs[k] = word & ~kFormedInner;
		if (e == kEmpty || word == kFormed || (word & kFormedInner)) continue;
		const uint32_t q = word - 1u, s = (uint32_t)(e >> 32);
		const int ds = rdnear[q];
		ClimbNode w;
		if (rleft[q]) { w.l = rfar[q]; w.r = s; w.dl = rdfar[q]; w.dr = ds; }
		else { w.l = s + 1u; w.r = rfar[q]; w.dl = ds; w.dr = rdfar[q]; }
		push(frontier_entry(w, leaf0));
	}
	uint2 links[kPer];
	#pragma unroll
	for (int k = 0; k < kPer; ++k)
		if (inner[k]) links[k] = *reinterpret_cast<const uint2 *>(&nodes[xs[k]]);
	rta_node nl[kPer], nr[kPer];
	#pragma unroll
	for (int k = 0; k < kPer; ++k)
		if (inner[k]) {
			nl[k] = nodes[links[k].x >> 1];
			nr[k] = nodes[links[k].y];
		}
	bool slow = false;
	#pragma unroll
	for (int k = 0; k < kPer; ++k) {
		if (!inner[k]) continue;
		const uint32_t L = links[k].x >> 1, R = links[k].y;
		const uint32_t lw = link_inner(nl[k].type_left_elems) ? (L << 1) | 1u : nl[k].right_tris << 1;
		const uint32_t rw = link_inner(nr[k].type_left_elems) ? (R << 1) | 1u : nr[k].right_tris << 1;
		if (!write_pair(pk.pairs, xs[k], nl[k].box, nr[k].box, lw, rw, L, R)) slow = true;
	}
	if (slow && pk.flags) atomicOr(pk.flags, 1u);
	__syncthreads();
	const uint32_t nf = min(nfront, (uint32_t)kSlots);
	if (out_slots && (uint32_t)t < nf) out_slots[t] = front[t];
	if (out_count && t == 0) *out_count = nf;
}

// -------------------------------------------------- stage 7: permute

// out[i] = scratch[indices[i]] (invert_triangle_permutation_cuda,
// lbvh-kernels.cu:557-578) for the kPermuteBlock triangles of one block,
// as 13 * kPermuteBlock consecutive 8-byte output words: word w comes from
// word w % 13 of triangle indices[block * kPermuteBlock + w / 13].  Every
// store of a wave is contiguous and each scattered source record is read
// by 13 neighbouring lanes.  All threads of the block must call it.
constexpr int kPermuteBlock = 256;
constexpr int kVertexWords = 5;   // the 36 bytes of a, b, c lie in words 0..4
__device__ inline void permute_block(rta_triangle *__restrict__ out,
                                     const rta_triangle *__restrict__ scratch,
                                     const uint32_t *__restrict__ indices, int n, int block,
                                     float4 *__restrict__ tri4 = nullptr) {
	__shared__ uint32_t src[kPermuteBlock];
	__shared__ uint2 verts[kPermuteBlock][kVertexWords];
	const int t = threadIdx.x;
	const int base = block * kPermuteBlock;
	const int count = min(kPermuteBlock, n - base);
	if (t < count) src[t] = indices[base + t];
	__syncthreads();
	const uint2 *s = reinterpret_cast<const uint2 *>(scratch);
	uint2 *d = reinterpret_cast<uint2 *>(out + base);
	const int nwords = count * kTriWords;
	uint2 v[kTriWords];
#pragma unroll
	for (int k = 0; k < kTriWords; ++k) {
		const int w = t + k * kPermuteBlock;
		if (w < nwords) {
			const int r = w / kTriWords;
			v[k] = s[(size_t)src[r] * kTriWords + (w - r * kTriWords)];
		}
	}
#pragma unroll
	for (int k = 0; k < kTriWords; ++k) {
		const int w = t + k * kPermuteBlock;
		if (w < nwords) d[w] = v[k];
	}
	if (tri4) {
		// the vertex words of each triangle, gathered in LDS, give its three
		// tri4 records (a, a - b, a - c; every triangle first marked as the
		// end of a leaf, the bottom-up pass clears the marks that collapses
		// make interior)
#pragma unroll
		for (int k = 0; k < kTriWords; ++k) {
			const int w = t + k * kPermuteBlock;
			if (w < nwords) {
				const int r = w / kTriWords, c = w - r * kTriWords;
				if (c < kVertexWords) verts[r][c] = v[k];
			}
		}
		__syncthreads();
		if (t < count) {
			const float *f = reinterpret_cast<const float *>(&verts[t][0]);
			const float ax = f[0], ay = f[1], az = f[2];
			float4 *o = tri4 + 3 * (size_t)(base + t);
			o[0] = make_float4(ax, ay, az, __uint_as_float(1u));
			o[1] = make_float4(ax - f[3], ay - f[4], az - f[5], 0.0f);
			o[2] = make_float4(ax - f[6], ay - f[7], az - f[8], 0.0f);
		}
	}
}

}  // namespace rta_hip
