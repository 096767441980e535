// Spatial-median binary BVH built on the host: what the bbvh-cuda plugin
// builds on the CPU and uploads for `-b median` (bbvh-cuda/main.cpp:
// 158-163, 200-235), i.e. bbvh_constructor_using_median with
// spatial_median (bbvh/bbvh.h:110-279).
//
// What the result is (bit-identical to the reference on finite input):
//  - a node's box is the AABB of the vertices of its range, started from the
//    first triangle's vertex a and updated on strict < / > vertex by vertex
//    (compute_aabb / merge, librta/aabb.h:10-28, 40-49);
//  - a range of more than 4 triangles is split: axis = x if dx > dy and
//    dx > dz, y if dy >= dx and dy > dz, z otherwise (bbvh.h:218-223); the
//    range is ordered by that component of the centroid,
//    float((a + b + c) * 0.33333) with the sum in float and the product in
//    double (librta/tri.h:8-14), keeping the previous order among equal
//    components (glibc's qsort is a merge sort for arrays it can buffer);
//    the split position is the first one whose component is not below
//    (max - min) / 2 + min of the range's components, or the middle when
//    that leaves a side empty (bbvh.h:187-204, 225-227);
//  - nodes are numbered in pre-order (a node, its left subtree, its right
//    subtree); a leaf's triangles follow the previous leaf's.
// NaN centroid components make the reference's comparator inconsistent;
// their order is not defined here either.
//
// How: an explicit work stack instead of recursion; each split orders its
// range with a stable LSD radix sort on an order-preserving 32-bit key of
// the component (8-bit digits, digits shared by the whole range skipped),
// or an insertion sort for short ranges.  O(n log n) over the tree.
#include <string.h>

#include <cstdint>
#include <vector>

#include "../../../include/rta_host.h"

namespace {

// Order-preserving key of a float: -0 and +0 map to one key.
inline uint32_t sort_key(float f) {
	if (f == 0.0f) f = 0.0f;
	uint32_t u;
	memcpy(&u, &f, 4);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct Task {
	uint32_t begin, end;
	uint32_t parent;   // node whose right link this subtree fills (~0u: none)
};

class MedianBuild {
public:
	MedianBuild(const rta_triangle *tris, uint32_t n) : tris_(tris), n_(n) {
		for (int ax = 0; ax < 3; ++ax) {
			key_[ax].resize(n);
			cog_[ax].resize(n);
		}
		for (uint32_t i = 0; i < n; ++i) {
			const float *a = &tris[i].a.x, *b = &tris[i].b.x, *c = &tris[i].c.x;
			for (int ax = 0; ax < 3; ++ax) {
				const float s = a[ax] + b[ax] + c[ax];
				const float g = (float)((double)s * 0.33333);
				cog_[ax][i] = g;
				key_[ax][i] = sort_key(g);
			}
		}
		perm_.resize(n);
		for (uint32_t i = 0; i < n; ++i) perm_[i] = i;
		tmp_.resize(n);
	}

	void run(std::vector<rta_node> &nodes, std::vector<uint32_t> &order) {
		nodes.clear();
		order.clear();
		nodes.reserve(2 * (size_t)n_);
		order.reserve(n_);
		std::vector<Task> stack;
		stack.push_back({0u, n_, ~0u});
		while (!stack.empty()) {
			const Task t = stack.back();
			stack.pop_back();
			const uint32_t id = (uint32_t)nodes.size();
			if (t.parent != ~0u) nodes[t.parent].right_tris = id;
			rta_node nd;
			memset(&nd, 0, sizeof nd);
			nd.box = range_box(t.begin, t.end);
			const uint32_t elems = t.end - t.begin;
			if (elems > kMaxLeaf) {
				const float dx = nd.box.max.x - nd.box.min.x;
				const float dy = nd.box.max.y - nd.box.min.y;
				const float dz = nd.box.max.z - nd.box.min.z;
				const int axis = dx > dy ? (dx > dz ? 0 : 2) : (dy > dz ? 1 : 2);
				sort_range(t.begin, t.end, axis);
				uint32_t mid = split_position(t.begin, t.end, axis);
				if (mid == t.begin || mid == t.end) mid = t.begin + elems / 2;
				// the left child is the next node; the right one fills in
				// its id when its task comes off the stack
				nd.type_left_elems = ((id + 1) << 1) | 1u;
				nodes.push_back(nd);
				stack.push_back({mid, t.end, id});
				stack.push_back({t.begin, mid, ~0u});
			} else {
				nd.type_left_elems = elems << 1;
				nd.right_tris = (uint32_t)order.size();
				nodes.push_back(nd);
				for (uint32_t i = t.begin; i < t.end; ++i) order.push_back(perm_[i]);
			}
		}
	}

private:
	static constexpr uint32_t kMaxLeaf = 4;     // default_max_tris_per_node, bbvh.h:252
	static constexpr uint32_t kInsertion = 32;  // ranges up to this use insertion sort

	rta_aabb range_box(uint32_t begin, uint32_t end) const {
		rta_aabb bb;
		bb.min = tris_[perm_[begin]].a;
		bb.max = bb.min;
		float *lo = &bb.min.x, *hi = &bb.max.x;
		for (uint32_t i = begin; i < end; ++i) {
			const rta_triangle &t = tris_[perm_[i]];
			for (const rta_float3 *v : {&t.a, &t.b, &t.c}) {
				const float *p = &v->x;
				for (int c = 0; c < 3; ++c) {
					if (p[c] < lo[c]) lo[c] = p[c];
					if (p[c] > hi[c]) hi[c] = p[c];
				}
			}
		}
		return bb;
	}

	// stable order of perm_[begin, end) by key_[axis]
	void sort_range(uint32_t begin, uint32_t end, int axis) {
		const uint32_t *key = key_[axis].data();
		uint32_t *p = perm_.data() + begin;
		const uint32_t m = end - begin;
		if (m <= kInsertion) {
			for (uint32_t i = 1; i < m; ++i) {
				const uint32_t v = p[i], k = key[v];
				uint32_t j = i;
				while (j > 0 && key[p[j - 1]] > k) {
					p[j] = p[j - 1];
					--j;
				}
				p[j] = v;
			}
			return;
		}
		uint32_t all_or = 0, all_and = ~0u;
		for (uint32_t i = 0; i < m; ++i) {
			all_or |= key[p[i]];
			all_and &= key[p[i]];
		}
		uint32_t *src = p, *dst = tmp_.data() + begin;
		for (int shift = 0; shift < 32; shift += 8) {
			if ((((all_or ^ all_and) >> shift) & 0xffu) == 0u) continue;   // one digit value
			uint32_t count[257] = {0};
			for (uint32_t i = 0; i < m; ++i) ++count[((key[src[i]] >> shift) & 0xffu) + 1];
			for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
			for (uint32_t i = 0; i < m; ++i) dst[count[(key[src[i]] >> shift) & 0xffu]++] = src[i];
			uint32_t *t = src;
			src = dst;
			dst = t;
		}
		if (src != p) memcpy(p, src, sizeof(uint32_t) * m);
	}

	// compute_spatial_median over the sorted range (bbvh.h:187-204)
	uint32_t split_position(uint32_t begin, uint32_t end, int axis) const {
		const float *c = cog_[axis].data();
		float lo = c[perm_[begin]], hi = lo;
		for (uint32_t i = begin; i < end; ++i) {
			const float v = c[perm_[i]];
			if (v < lo) lo = v;
			if (v > hi) hi = v;
		}
		const float s = (hi - lo) / 2 + lo;
		uint32_t i = begin;
		while (i != end && c[perm_[i]] < s) ++i;
		return i;
	}

	const rta_triangle *tris_;
	uint32_t n_;
	std::vector<uint32_t> key_[3];
	std::vector<float> cog_[3];
	std::vector<uint32_t> perm_, tmp_;
};

}  // namespace

extern "C" {

int rta_host_median_node_bound(int n) { return n > 0 ? 2 * n - 1 : -1; }

int rta_host_median_build(const rta_triangle *tris, int n, rta_node *nodes, rta_triangle *tris_out,
                          uint32_t *order) {
	if (!tris || n <= 0 || !nodes) return -1;
	MedianBuild b(tris, (uint32_t)n);
	std::vector<rta_node> nv;
	std::vector<uint32_t> ov;
	b.run(nv, ov);
	memcpy(nodes, nv.data(), sizeof(rta_node) * nv.size());
	for (int i = 0; i < n; ++i) {
		if (tris_out) tris_out[i] = tris[ov[i]];
		if (order) order[i] = ov[i];
	}
	return (int)nv.size();
}

}  // extern "C"
