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
constexpr int kMergeFan = 64;
constexpr uint32_t kStateNone = 0u;           // lelems: node not completed
constexpr uint32_t kStateInner = 0xffffffffu; // lelems: completed inner node

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

// Bottom-up boxes (k::build_bounding_boxes_karras2, lbvh-kernels.cu:258-341):
// every leaf climbs towards the root; at each inner node the first arriving
// child stops and the second forms the union and continues.  Deterministic
// form of the reference:
//  * a separate visit counter decides first/second arrival (the reference
//    sums representative codes into the flag, so a first arrival with code
//    0 also stops the second thread, lbvh-kernels.cu:288-291);
//  * the union is MIN(left, right) / MAX(left, right) with the reference's
//    `(a<b)?a:b`, fixing the sign of zero that the reference leaves to
//    arrival order;
//  * merge rule: both children are leaves and their representative codes
//    (code of the leaf, or codes[i] for inner node i, :261) are equal
//    (`code == old` in the reference) -> the parent becomes one leaf with
//    tris = min, elems = sum (lbvh-kernels.cu:302-330).
//
// Hierarchical schedule.  A subtree covers a contiguous range of sorted
// positions, so all communication inside a chunk of consecutive leaves
// stays in one workgroup:
//  level 1: a 1024-thread block owns 1024 leaves; counters, boxes and leaf
//           state live in LDS; a climb stops when the parent's range leaves
//           the chunk and the child is recorded in the block's frontier
//           slots (at most 2 per tree level, so <= 126 for depth <= 63);
//  level k: one block merges the frontiers of kMergeFan level-(k-1) chunks
//           with workgroup-scope atomics on the global visit counters,
//           until one block covers all leaves and reaches the root.
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

// a collapse merges two leaves: the left one's last triangle no longer ends
// a leaf (the permute wrote every triangle as a leaf end)
__device__ inline void clear_leaf_end(float4 *tri4, uint32_t pos) {
	tri4[3 * (size_t)pos].w = 0.0f;
}

// one whole node: its link words and the box held in LDS column q of
// b[6][kChunk], as two 16-byte stores
__device__ inline void store_node(rta_node *node, uint2 links, const float (*b)[kChunk], int q) {
	uint4 *o = reinterpret_cast<uint4 *>(node);
	o[0] = make_uint4(links.x, links.y, __float_as_uint(b[0][q]), __float_as_uint(b[1][q]));
	o[1] = make_uint4(__float_as_uint(b[2][q]), __float_as_uint(b[3][q]), __float_as_uint(b[4][q]),
	                  __float_as_uint(b[5][q]));
}

__device__ inline uint32_t rep_code(const uint32_t *codes, uint32_t node, uint32_t leaf0) {
	return node >= leaf0 ? codes[node - leaf0] : codes[node];
}

// Level 1 for leaves [chunk*kChunk, +kChunk) of a mesh; the frontier goes
// to slots[0..kSlots) and its size to *count.
//
// An inner node whose range lies inside the chunk has its index inside it
// too (Karras: a node's index is one end of its range), and so do its
// children.  The block therefore first loads, coalesced, everything the
// climbs read: representative codes, each such node's children as 16-bit
// local references (kLocalLeaf marks a leaf), and each node's parent as a
// local index, kParentExit when the parent's range leaves the chunk (the
// climb stops and records a frontier entry) or kParentRoot for the root.
// The climbs then touch LDS only.
constexpr uint16_t kLocalLeaf = 0x8000u;
constexpr uint16_t kParentExit = 0xffffu;
constexpr uint16_t kParentRoot = 0xfffeu;

__device__ inline void bottom_up_chunk(rta_node *__restrict__ nodes,
                                       const rta_aabb *__restrict__ boxes,
                                       const uint32_t *__restrict__ codes,
                                       const uint32_t *__restrict__ indices,
                                       const uint32_t *__restrict__ parents,
                                       const uint2 *__restrict__ ranges,
                                       uint32_t *__restrict__ slots, uint32_t *__restrict__ count,
                                       uint32_t *__restrict__ status, int n, int chunk,
                                       PackOut pk = PackOut{}) {
	__shared__ float lb_leaf[6][kChunk];
	__shared__ float lb_int[6][kChunk];
	__shared__ uint32_t lvisit[kChunk];
	__shared__ uint32_t ltris[kChunk];
	__shared__ uint32_t lelems[kChunk];
	__shared__ uint32_t lcode[kChunk];
	__shared__ uint32_t lkids[kChunk];      // left | right << 16, ~0u: range leaves the chunk
	__shared__ uint16_t lpar_leaf[kChunk];
	__shared__ uint16_t lpar_int[kChunk];
	__shared__ uint32_t front[kSlots];
	__shared__ uint32_t nfront;

	const int t = threadIdx.x;
	const uint32_t base = (uint32_t)chunk * kChunk;
	const uint32_t leaf0 = (uint32_t)(n - 1);
	const uint32_t end = min(base + (uint32_t)kChunk, (uint32_t)n);
	const uint32_t k = base + t;          // leaf k and inner node k
	lvisit[t] = 0;
	lelems[t] = kStateNone;
	if (t == 0) nfront = 0;
	rta_aabb box;
	if (k < (uint32_t)n) {
		box = boxes[indices[k]];
		lb_leaf[0][t] = box.min.x; lb_leaf[1][t] = box.min.y; lb_leaf[2][t] = box.min.z;
		lb_leaf[3][t] = box.max.x; lb_leaf[4][t] = box.max.y; lb_leaf[5][t] = box.max.z;
		lcode[t] = codes[k];
	}
	const bool inner = k < leaf0 && k < end;
	uint2 own_links = make_uint2(0u, 0u);   // node k's links when its range is in the chunk
	if (inner) {
		const uint2 rg = ranges[k];
		uint32_t kids = ~0u;
		if (rg.x >= base && rg.y < end) {
			const uint2 pl = *reinterpret_cast<const uint2 *>(&nodes[k]);
			own_links = pl;
			const uint32_t l = pl.x >> 1, r = pl.y;
			const uint32_t ll = l >= leaf0 ? (l - leaf0 - base) | kLocalLeaf : l - base;
			const uint32_t rl = r >= leaf0 ? (r - leaf0 - base) | kLocalLeaf : r - base;
			kids = ll | (rl << 16);
		}
		lkids[t] = kids;
	}
	// parents are loaded with the rest; only the decision needs lkids
	const uint32_t leaf_parent = k < (uint32_t)n && leaf0 + k != 0u ? parents[leaf0 + k] : 0u;
	const uint32_t inner_parent = inner && k != 0u ? parents[k] : 0u;
	__syncthreads();
	auto local_parent = [&](uint32_t node, uint32_t p) -> uint16_t {
		if (node == 0u) return kParentRoot;
		if (p < base || p >= end || p >= leaf0 || lkids[p - base] == ~0u) return kParentExit;
		return (uint16_t)(p - base);
	};
	if (k < (uint32_t)n) lpar_leaf[t] = local_parent(leaf0 + k, leaf_parent);
	if (inner) lpar_int[t] = local_parent(k, inner_parent);
	__syncthreads();

	if (k < (uint32_t)n) {
		uint32_t cur = (uint32_t)t | kLocalLeaf;
		uint32_t p = lpar_leaf[t];
		for (int level = 0; level < 128; ++level) {
			if (p == kParentRoot) break;
			if (p == kParentExit) {
				const uint32_t g = (cur & kLocalLeaf) ? leaf0 + base + (cur & 0x3ffu) : base + cur;
				const uint32_t slot = atomicAdd(&nfront, 1u);
				if (slot < (uint32_t)kSlots) front[slot] = g;
				else atomicOr(status, (uint32_t)RTA_LBVH_FRONTIER_OVERFLOW);
				break;
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
			const uint32_t old = atomicAdd(&lvisit[p], 1u);
			if (old == 0u) break;
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");

			const uint32_t kids = lkids[p];
			const uint32_t left = kids & 0xffffu, right = kids >> 16;
			const bool cur_is_left = (left == cur);
			const uint32_t sib = cur_is_left ? right : left;
			const uint32_t q = sib & 0x3ffu;
			rta_aabb sbox;
			if (sib & kLocalLeaf) {
				sbox.min = {lb_leaf[0][q], lb_leaf[1][q], lb_leaf[2][q]};
				sbox.max = {lb_leaf[3][q], lb_leaf[4][q], lb_leaf[5][q]};
			} else {
				sbox.min = {lb_int[0][q], lb_int[1][q], lb_int[2][q]};
				sbox.max = {lb_int[3][q], lb_int[4][q], lb_int[5][q]};
			}
			box = cur_is_left ? union_lr(box, sbox) : union_lr(sbox, box);
			lb_int[0][p] = box.min.x; lb_int[1][p] = box.min.y; lb_int[2][p] = box.min.z;
			lb_int[3][p] = box.max.x; lb_int[4][p] = box.max.y; lb_int[5][p] = box.max.z;

			// children state: original leaf, collapsed inner node, or inner.
			// The representative code of leaf j and of inner node j are both
			// codes[j] (lbvh-kernels.cu:261), i.e. lcode of the local index.
			uint32_t state = kStateInner, ptris = 0;
			if (lcode[cur & 0x3ffu] == lcode[q]) {
				uint32_t le, lt, re, rt;
				bool l_leaf, r_leaf;
				const uint32_t li = left & 0x3ffu, ri = right & 0x3ffu;
				if (left & kLocalLeaf) { l_leaf = true; le = 1; lt = base + li; }
				else { le = lelems[li]; lt = ltris[li]; l_leaf = le != kStateInner; }
				if (right & kLocalLeaf) { r_leaf = true; re = 1; rt = base + ri; }
				else { re = lelems[ri]; rt = ltris[ri]; r_leaf = re != kStateInner; }
				if (l_leaf && r_leaf) { state = re + le; ptris = lt < rt ? lt : rt; }
			}
			ltris[p] = ptris;
			lelems[p] = state;
			cur = p;
			p = lpar_int[p];
		}
	}
	__syncthreads();

	// write-out: whole 32-byte leaf nodes (links: one triangle, the first
	// at position k) and completed inner nodes, then the frontier
	if (k < (uint32_t)n) store_node(&nodes[leaf0 + k], make_uint2(1u << 1, k), lb_leaf, t);
	bool slow = false;
	if (k < leaf0 && lelems[t] != kStateNone) {
		const bool collapsed = lelems[t] != kStateInner;
		store_node(&nodes[k], collapsed ? make_uint2(lelems[t] << 1, ltris[t]) : own_links, lb_int,
		           t);
		if (collapsed) {
			// the collapse merged two leaves that meet at the node's split,
			// the left child's last position: base + its local reference
			// (a leaf's position, or an inner left child's index, which is
			// the end of its range)
			if (pk.tri4) clear_leaf_end(pk.tri4, base + (lkids[t] & 0x3ffu));
		} else if (pk.pairs) {
			// a completed inner node: its pair record from the children's
			// final state (an original leaf, a collapsed node or an inner one)
			const uint32_t kids = lkids[t];
			rta_aabb cb[2];
			uint32_t word[2], id[2];
			for (int c = 0; c < 2; ++c) {
				const uint32_t ref = c == 0 ? (kids & 0xffffu) : (kids >> 16);
				const uint32_t q = ref & 0x3ffu;
				const float(*src)[kChunk] = (ref & kLocalLeaf) ? lb_leaf : lb_int;
				cb[c].min = {src[0][q], src[1][q], src[2][q]};
				cb[c].max = {src[3][q], src[4][q], src[5][q]};
				if (ref & kLocalLeaf) {
					word[c] = (base + q) << 1;
					id[c] = leaf0 + base + q;
				} else {
					word[c] = lelems[q] == kStateInner ? ((base + q) << 1) | 1u : ltris[q] << 1;
					id[c] = base + q;
				}
			}
			slow = !write_pair(pk.pairs, k, cb[0], cb[1], word[0], word[1], id[0], id[1]);
		}
	}
	if (pk.flags && __syncthreads_or(slow) && t == 0) atomicOr(pk.flags, 1u);
	const uint32_t nf = min(nfront, (uint32_t)kSlots);
	if ((uint32_t)t < nf) slots[t] = front[t];
	if (t == 0) *count = nf;
}

__device__ inline rta_aabb load_box(const rta_node *node) {
	const float *f = reinterpret_cast<const float *>(node) + 2;
	rta_aabb b;
	b.min = {f[0], f[1], f[2]};
	b.max = {f[3], f[4], f[5]};
	return b;
}
__device__ inline void store_box(rta_node *node, const rta_aabb &b) {
	float *f = reinterpret_cast<float *>(node) + 2;
	f[0] = b.min.x; f[1] = b.min.y; f[2] = b.min.z;
	f[3] = b.max.x; f[4] = b.max.y; f[5] = b.max.z;
}

// Level k >= 2 for one mesh: merges the frontier slots of the ng chunks of
// the previous level starting at in_slots/in_counts (kSlots words per
// chunk), whose leaves span [lo, hi); the new frontier goes to out_slots /
// *out_count when given.
__device__ inline void bottom_up_merge(rta_node *__restrict__ nodes, uint32_t *__restrict__ visit,
                                       const uint32_t *__restrict__ codes,
                                       const uint32_t *__restrict__ parents,
                                       const uint2 *__restrict__ ranges,
                                       const uint32_t *__restrict__ in_slots,
                                       const uint32_t *__restrict__ in_counts, int ng,
                                       uint32_t *__restrict__ out_slots,
                                       uint32_t *__restrict__ out_count,
                                       uint32_t *__restrict__ status, uint32_t lo, uint32_t hi,
                                       int n, PackOut pk = PackOut{}) {
	__shared__ uint32_t prefix[kMergeFan + 1];
	__shared__ uint32_t front[kSlots];
	__shared__ uint32_t nfront;
	const int t = threadIdx.x;
	if (t == 0) {
		uint32_t acc = 0;
		prefix[0] = 0;
		for (int g = 0; g < ng; ++g) {
			acc += in_counts[g];
			prefix[g + 1] = acc;
		}
		nfront = 0;
	}
	__syncthreads();
	const uint32_t leaf0 = (uint32_t)(n - 1);
	const uint32_t total = prefix[ng];
	for (uint32_t f = t; f < total; f += blockDim.x) {
		int g = 0;
		while (prefix[g + 1] <= f) ++g;
		uint32_t cur = in_slots[(size_t)g * kSlots + (f - prefix[g])];
		if (cur == 0u) continue;
		rta_aabb box = load_box(&nodes[cur]);
		uint32_t code_cur = rep_code(codes, cur, leaf0);
		uint32_t parent = parents[cur];
		// Per level: (a) what earlier stages wrote and nothing changes
		// before this thread arrives -- a node's range, links, code and
		// parent; (b) the visit counter; (c) after its acquire, what the
		// sibling subtree's thread wrote -- the sibling's box and the
		// children's (possibly collapsed) links.  (a) for the grandparent is
		// issued before (b), so a level costs two dependent round trips.
		// The grandparent's links change only when a thread arrives there
		// second; this thread reads them before it could arrive, and uses
		// them only if it does.
		uint2 rg = ranges[parent];
		uint2 pl = *reinterpret_cast<const uint2 *>(&nodes[parent]);
		uint32_t grand = parent != 0u ? parents[parent] : 0u;
		uint32_t code_parent = codes[parent];
		for (int level = 0; cur != 0u && level < 128; ++level) {
			if (rg.x < lo || rg.y >= hi) {
				const uint32_t slot = atomicAdd(&nfront, 1u);
				if (slot < (uint32_t)kSlots) front[slot] = cur;
				else atomicOr(status, (uint32_t)RTA_LBVH_FRONTIER_OVERFLOW);
				break;
			}
			uint2 rg_next = make_uint2(0u, 0u), pl_next = make_uint2(0u, 0u);
			uint32_t grand_next = 0u, code_next = 0u;
			if (parent != 0u) {
				rg_next = ranges[grand];
				pl_next = *reinterpret_cast<const uint2 *>(&nodes[grand]);
				grand_next = grand != 0u ? parents[grand] : 0u;
				code_next = codes[grand];
			}
			// acquire only: this thread's earlier box and link stores were
			// released by the fence that ends the previous level (or by the
			// launch boundary), so the prefetch above need not complete first
			const uint32_t old = __hip_atomic_fetch_add(&visit[parent], 1u, __ATOMIC_ACQUIRE,
			                                            __HIP_MEMORY_SCOPE_WORKGROUP);
			if (old == 0u) break;
			const uint32_t left = pl.x >> 1, right = pl.y;
			const bool cur_is_left = (left == cur);
			const uint32_t sib = cur_is_left ? right : left;
			const rta_aabb sbox = load_box(&nodes[sib]);
			const uint32_t code_sib = rep_code(codes, sib, leaf0);
			const uint2 ll = *reinterpret_cast<const uint2 *>(&nodes[left]);
			const uint2 rl = *reinterpret_cast<const uint2 *>(&nodes[right]);
			const rta_aabb lbox = cur_is_left ? box : sbox, rbox = cur_is_left ? sbox : box;
			box = union_lr(lbox, rbox);
			store_box(&nodes[parent], box);
			const bool collapse = code_cur == code_sib && !link_inner(ll.x) && !link_inner(rl.x);
			if (collapse) {
				const uint32_t tris = ll.y < rl.y ? ll.y : rl.y;
				const uint32_t elems = (rl.x >> 1) + (ll.x >> 1);
				*reinterpret_cast<uint2 *>(&nodes[parent]) = make_uint2(elems << 1, tris);
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
			// the leaf-end mark is read only after the build: its store
			// follows the fence, which waits for the box and links only.
			// (Pair records of the nodes completed here are written by
			// pack_boundary_pairs after the last level.)
			if (collapse && pk.tri4) clear_leaf_end(pk.tri4, ll.y + (ll.x >> 1) - 1u);
			cur = parent;
			parent = grand;
			code_cur = code_parent;
			rg = rg_next;
			pl = pl_next;
			grand = grand_next;
			code_parent = code_next;
		}
	}
	__syncthreads();
	const uint32_t nf = min(nfront, (uint32_t)kSlots);
	if (out_slots && (uint32_t)t < nf) out_slots[t] = front[t];
	if (out_count && t == 0) *out_count = nf;
}

// Pair records of the inner nodes the chunk level cannot complete: those
// whose range [from, to] spans more than one kChunk-leaf chunk, finished by
// the merge levels.  Node x is one of them iff from / kChunk != to / kChunk;
// its record comes from the finished nodes, as rta_bvh_pack forms it.
__device__ inline void pack_boundary_pair(rta_node_pair *__restrict__ pairs,
                                          uint32_t *__restrict__ flags,
                                          const rta_node *__restrict__ nodes,
                                          const uint2 *__restrict__ ranges, uint32_t x) {
	const uint2 rg = ranges[x];
	if (rg.x / (uint32_t)kChunk == rg.y / (uint32_t)kChunk) return;
	const uint2 links = *reinterpret_cast<const uint2 *>(&nodes[x]);
	if (!link_inner(links.x)) return;   // collapsed into a leaf
	const uint32_t L = links.x >> 1, R = links.y;
	const rta_node nl = nodes[L], nr = nodes[R];
	const uint32_t lw = link_inner(nl.type_left_elems) ? (L << 1) | 1u : nl.right_tris << 1;
	const uint32_t rw = link_inner(nr.type_left_elems) ? (R << 1) | 1u : nr.right_tris << 1;
	if (!write_pair(pairs, x, nl.box, nr.box, lw, rw, L, R)) atomicOr(flags, 1u);
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
