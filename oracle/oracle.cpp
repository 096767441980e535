// ORACLE -- test infrastructure only.  A CPU restatement of the reference's
// algorithms for the GPU BVH path, used by the tests, by smoke() and by the
// benchmark's CPU-baseline leg as the checker and the CPU baseline; nothing
// in the product links, loads or calls it.
//
// The reference itself (kiselgra/rta) cannot be compiled here: its headers
// need libmcm/vectors.h, png/png++ and a configure-generated rta-config.h,
// none of which exist in this image, so every function below restates the
// cited reference lines in plain C++.  Build: oracle/Makefile with
// -ffp-contract=off (no FMA contraction, SSE binary32 arithmetic), which is
// how the reference's CPU plugin computes.
//
// Parity pins (tests/test_oracle_*.py): Morton known answers and a bit-loop
// reference; numpy stable argsort; a brute-force binary radix tree over the
// augmented (code, position) keys; the range characterisation of boxes and
// collapsed leaves; exact-rational triangle/box cases and float64 Cramer;
// brute-force closest hit on small scenes; LBVH vs. median-tree agreement.
#include <stdint.h>
#include <float.h>
#include <math.h>
#include <string.h>
#include <sys/time.h>

#include <algorithm>
#include <vector>

#include <omp.h>

#include "../include/rta_hip.h"

namespace {

// ---------------------------------------------------------------- tests

// intersect_tri_opt, librta/intersect.h:8-53
bool intersect_tri_opt(const rta_triangle &t, const float *o, const float *dir, float &t_out,
                       float &beta_out, float &gamma_out) {
	float a_x = t.a.x;
	float a_y = t.a.y;
	float a_z = t.a.z;

	const float a = a_x - t.b.x;
	const float b = a_y - t.b.y;
	const float c = a_z - t.b.z;
	const float d = a_x - t.c.x;
	const float e = a_y - t.c.y;
	const float f = a_z - t.c.z;
	const float g = dir[0];
	const float h = dir[1];
	const float i = dir[2];
	const float j = a_x - o[0];
	const float k = a_y - o[1];
	const float l = a_z - o[2];

	float common1 = e * i - h * f;
	float common2 = g * f - d * i;
	float common3 = d * h - e * g;
	float M = a * common1 + b * common2 + c * common3;
	float beta = j * common1 + k * common2 + l * common3;

	common1 = a * k - j * b;
	common2 = j * c - a * l;
	common3 = b * l - k * c;
	float gamma = i * common1 + h * common2 + g * common3;
	float tt = -(f * common1 + e * common2 + d * common3);

	beta /= M;
	gamma /= M;
	tt /= M;

	if (tt > 0)
		if (beta > 0 && gamma > 0 && beta + gamma <= 1) {
			t_out = tt;
			beta_out = beta;
			gamma_out = gamma;
			return true;
		}
	return false;
}

// intersect_aabb, librta/intersect.h:56-150 (sequential, with early outs)
bool intersect_aabb(const rta_aabb &box, const float *o, const float *dir, float &is) {
	float t_near = -FLT_MAX;
	float t_far = FLT_MAX;
	const float mn[3] = {box.min.x, box.min.y, box.min.z};
	const float mx[3] = {box.max.x, box.max.y, box.max.z};
	for (int ax = 0; ax < 3; ++ax) {
		const float d_ = dir[ax];
		const float e_ = o[ax];
		const float x_max = mx[ax];
		const float x_min = mn[ax];
		if (d_ == 0) {
			if (e_ < x_min || e_ > x_max) return false;
		} else {
			float t1 = (x_min - e_) / d_;
			float t2 = (x_max - e_) / d_;
			if (t1 > t2) { float tmp = t1; t1 = t2; t2 = tmp; }
			if (t1 > t_near) t_near = t1;
			if (t2 < t_far) t_far = t2;
			if (t_near > t_far) return false;
			if (t_far < 0) return false;
		}
	}
	is = t_near;
	return true;
}

// ---------------------------------------------------------- LBVH pieces

// MIN/MAX macros of lbvh-kernels.cu:142-143,259-260,498-499
inline float MINF(float a, float b) { return (a < b) ? a : b; }
inline float MAXF(float a, float b) { return (a > b) ? a : b; }

// Morton_3D_Encode_10bit, lbvh-kernels.cu:99-129
uint32_t morton_3d_encode_10bit(uint32_t index1, uint32_t index2, uint32_t index3) {
	index1 &= 0x000003ff;
	index2 &= 0x000003ff;
	index3 &= 0x000003ff;
	index1 |= (index1 << 16);
	index2 |= (index2 << 16);
	index3 |= (index3 << 16);
	index1 &= 0x030000ff;
	index2 &= 0x030000ff;
	index3 &= 0x030000ff;
	index1 |= (index1 << 8);
	index2 |= (index2 << 8);
	index3 |= (index3 << 8);
	index1 &= 0x0300f00f;
	index2 &= 0x0300f00f;
	index3 &= 0x0300f00f;
	index1 |= (index1 << 4);
	index2 |= (index2 << 4);
	index3 |= (index3 << 4);
	index1 &= 0x030c30c3;
	index2 &= 0x030c30c3;
	index3 &= 0x030c30c3;
	index1 |= (index1 << 2);
	index2 |= (index2 << 2);
	index3 |= (index3 << 2);
	index1 &= 0x09249249;
	index2 &= 0x09249249;
	index3 &= 0x09249249;
	return (index1 | (index2 << 1) | (index3 << 2));
}

// the reference's implicit float->uint conversion `uint x = center.x*d.x`
// (lbvh-kernels.cu:177) compiled by nvcc: cvt.rzi.u32.f32, i.e. truncation
// with saturation and NaN -> 0
uint32_t cvt_rzi_u32(float f) {
	if (isnan(f) || f <= 0.0f) return 0u;
	if (f >= 4294967296.0f) return 0xffffffffu;
	return (uint32_t)f;
}

// __cpl, lbvh-kernels.cu:84-92
int cpl(int m1, int m2, const uint32_t *codes, int n) {
	if (m2 < 0 || m2 >= n) return -1;
	if (codes[m1] == codes[m2]) return __builtin_clz((uint32_t)(m1 ^ m2)) + 32;
	return __builtin_clz(codes[m1] ^ codes[m2]);
}

inline void set_links(rta_node &nd, uint32_t tle, uint32_t rt) {
	nd.type_left_elems = tle;
	nd.right_tris = rt;
}

double now_ms() {
	timeval tv;
	gettimeofday(&tv, 0);
	return (double)tv.tv_sec * 1000.0 + (double)tv.tv_usec / 1000.0;
}

// ------------------------------------------------------ median ctor

// center_of_gravity, librta/tri.h:8-14: float sum, times the double 0.33333,
// stored back to float
inline float cog_comp(const rta_triangle &t, int axis) {
	const float *a = &t.a.x, *b = &t.b.x, *c = &t.c.x;
	return (float)((a[axis] + b[axis] + c[axis]) * 0.33333);
}

// merge / compute_aabb, librta/aabb.h:10-28,40-49
void merge_tri(rta_aabb &bb, const rta_triangle &t) {
	const rta_float3 *v[3] = {&t.a, &t.b, &t.c};
	for (int k = 0; k < 3; ++k) {
		const rta_float3 &p = *v[k];
		if (p.x < bb.min.x) bb.min.x = p.x;
		if (p.x > bb.max.x) bb.max.x = p.x;
		if (p.y < bb.min.y) bb.min.y = p.y;
		if (p.y > bb.max.y) bb.max.y = p.y;
		if (p.z < bb.min.z) bb.min.z = p.z;
		if (p.z > bb.max.z) bb.max.z = p.z;
	}
}

struct MedianBuilder {
	const rta_triangle *in;
	std::vector<uint32_t> perm;         // current order of input triangles
	std::vector<float> cog[3];          // centroid components per input tri
	std::vector<rta_node> nodes;
	std::vector<uint32_t> out_order;    // triangles in leaf order
	uint32_t max_tris_per_node = 4;     // bbvh.h:252

	rta_aabb compute_aabb(uint32_t begin, uint32_t end) {
		rta_aabb bb;
		bb.min = in[perm[begin]].a;
		bb.max = in[perm[begin]].a;
		for (uint32_t it = begin; it != end; ++it) merge_tri(bb, in[perm[it]]);
		return bb;
	}

	// compute_spatial_median, bbvh/bbvh.h:187-204
	uint32_t spatial_median(uint32_t begin, uint32_t end, int axis) {
		const std::vector<float> &c = cog[axis];
		float min_cog = c[perm[begin]], max_cog = c[perm[begin]];
		for (uint32_t it = begin; it != end; ++it) {
			float tmp = c[perm[it]];
			if (tmp < min_cog) min_cog = tmp;
			if (tmp > max_cog) max_cog = tmp;
		}
		float s = (max_cog - min_cog) / 2 + min_cog;
		uint32_t it = begin;
		while (it != end && c[perm[it]] < s) ++it;
		return it;
	}

	// build_sm, bbvh/bbvh.h:206-248.  glibc 2.35's qsort is a stable merge
	// sort at these sizes, so the per-node qsort by one centroid component
	// is std::stable_sort by that component.
	uint32_t build_sm(uint32_t begin, uint32_t end) {
		uint32_t id = (uint32_t)nodes.size();
		nodes.push_back(rta_node());
		memset(&nodes[id], 0, sizeof(rta_node));
		rta_aabb box = compute_aabb(begin, end);
		nodes[id].box = box;
		uint32_t elems = end - begin;
		if (elems > max_tris_per_node) {
			float dx = box.max.x - box.min.x, dy = box.max.y - box.min.y,
			      dz = box.max.z - box.min.z;
			int axis;
			if (dx > dy)
				axis = (dx > dz) ? 0 : 2;
			else
				axis = (dy > dz) ? 1 : 2;
			const std::vector<float> &c = cog[axis];
			std::stable_sort(perm.begin() + begin, perm.begin() + end,
			                 [&](uint32_t p, uint32_t q) { return c[p] < c[q]; });
			uint32_t mid = spatial_median(begin, end, axis);
			if (begin == mid || mid == end) mid = begin + (end - begin) / 2;
			uint32_t left = build_sm(begin, mid);
			uint32_t right = build_sm(mid, end);
			set_links(nodes[id], (left << 1) | 1u, right);
		} else {
			uint32_t idx = (uint32_t)out_order.size();
			for (uint32_t i = 0; i < elems; ++i) out_order.push_back(perm[begin + i]);
			set_links(nodes[id], elems << 1, idx);
		}
		return id;
	}
};

}  // namespace

extern "C" {

// ---------------------------------------------------------- single tests

int oracle_intersect_tri(const rta_triangle *t, const float *o, const float *d, float *out3) {
	float tt = 0, be = 0, ga = 0;
	bool hit = intersect_tri_opt(*t, o, d, tt, be, ga);
	if (hit) { out3[0] = tt; out3[1] = be; out3[2] = ga; }
	return hit ? 1 : 0;
}

int oracle_intersect_aabb(const rta_aabb *box, const float *o, const float *d, float *dist) {
	float is = 0;
	bool hit = intersect_aabb(*box, o, d, is);
	if (hit) *dist = is;
	return hit ? 1 : 0;
}

uint32_t oracle_morton_encode(uint32_t x, uint32_t y, uint32_t z) {
	return morton_3d_encode_10bit(x, y, z);
}

// --------------------------------------------------------------- LBVH

// lbvh_constructor::buildCUDA (bbvh-cuda/lbvh.h:153-211) with each stage of
// bbvh-cuda/lbvh-kernels.cu restated sequentially.  Outputs:
//   tris_out[n]          AS triangle array (Morton order)
//   nodes[2n-1]          32-byte nodes
//   indices[n]           sorted position -> input index
//   codes_sorted[n]      sorted Morton codes
//   codes_input[n]       Morton code of each input triangle (may be NULL)
//   parents[2n-1]        parent of each node (parents[0] is never written
//                        by the reference; set to 0xffffffff here)
//   bounds               scene box (may be NULL)
// Returns the status bits (RTA_LBVH_CODE_ZERO), or -1 for n <= 0.
int oracle_lbvh_build(const rta_triangle *tris_in, int n, rta_triangle *tris_out,
                      rta_node *nodes, uint32_t *indices, uint32_t *codes_sorted,
                      uint32_t *codes_input, uint32_t *parents, rta_aabb *bounds) {
	if (n <= 0) return -1;
	int status = 0;
	// initTriangles (lbvh-kernels.cu:344-359)
	std::vector<rta_triangle> scratch(tris_in, tris_in + n);
	std::vector<rta_aabb> boxes(n);
	std::vector<rta_float3> centers(n);
	for (int id = 0; id < n; ++id) {
		const rta_triangle &t = scratch[id];
		rta_aabb &b = boxes[id];
		b.min.x = MINF(MINF(t.a.x, t.b.x), t.c.x);
		b.min.y = MINF(MINF(t.a.y, t.b.y), t.c.y);
		b.min.z = MINF(MINF(t.a.z, t.b.z), t.c.z);
		b.max.x = MAXF(MAXF(t.a.x, t.b.x), t.c.x);
		b.max.y = MAXF(MAXF(t.a.y, t.b.y), t.c.y);
		b.max.z = MAXF(MAXF(t.a.z, t.b.z), t.c.z);
		// center_of_gravity, lbvh-kernels.cu:60-63: (tri.a+tri.b+tri.c)/3.0f
		centers[id].x = ((t.a.x + t.b.x) + t.c.x) / 3.0f;
		centers[id].y = ((t.a.y + t.b.y) + t.c.y) / 3.0f;
		centers[id].z = ((t.a.z + t.b.z) + t.c.z) / 3.0f;
	}
	// calculateVoxelBounds (lbvh-kernels.cu:518-535), reduce_max (:496-516)
	// A NaN box coordinate is skipped: with a NaN operand MIN/MAX is neither
	// associative nor commutative, so the reference's thrust::reduce result
	// depends on its reduction order (parity unpinned there); without NaNs
	// this is the reference's min/max.
	rta_aabb vb;
	vb.min = {100000000000.0f, 100000000000.0f, 100000000000.0f};
	vb.max = {-100000000000.0f, -100000000000.0f, -100000000000.0f};
	auto fold_min = [](float acc, float v) { return v == v ? MINF(acc, v) : acc; };
	auto fold_max = [](float acc, float v) { return v == v ? MAXF(acc, v) : acc; };
	for (int i = 0; i < n; ++i) {
		vb.min.x = fold_min(vb.min.x, boxes[i].min.x);
		vb.min.y = fold_min(vb.min.y, boxes[i].min.y);
		vb.min.z = fold_min(vb.min.z, boxes[i].min.z);
		vb.max.x = fold_max(vb.max.x, boxes[i].max.x);
		vb.max.y = fold_max(vb.max.y, boxes[i].max.y);
		vb.max.z = fold_max(vb.max.z, boxes[i].max.z);
	}
	if (bounds) *bounds = vb;
	// calculateMortonCodes2 (lbvh-kernels.cu:405-421) + morton_codes2 (:169-183)
	float dx = 1024.0f / (vb.max.x - vb.min.x);
	float dy = 1024.0f / (vb.max.y - vb.min.y);
	float dz = 1024.0f / (vb.max.z - vb.min.z);
	std::vector<uint32_t> codes(n);
	for (int tid = 0; tid < n; ++tid) {
		float cx = centers[tid].x - vb.min.x;
		float cy = centers[tid].y - vb.min.y;
		float cz = centers[tid].z - vb.min.z;
		uint32_t x = cvt_rzi_u32(cx * dx);
		uint32_t y = cvt_rzi_u32(cy * dy);
		uint32_t z = cvt_rzi_u32(cz * dz);
		codes[tid] = morton_3d_encode_10bit(x, y, z);
		if (codes[tid] == 0) status |= RTA_LBVH_CODE_ZERO;
	}
	if (codes_input) memcpy(codes_input, codes.data(), sizeof(uint32_t) * n);
	// sortMortonCodes (lbvh-kernels.cu:537-555): stable sort by key
	std::vector<uint32_t> idx(n);
	for (int i = 0; i < n; ++i) idx[i] = (uint32_t)i;
	std::stable_sort(idx.begin(), idx.end(),
	                 [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
	std::vector<uint32_t> sc(n);
	for (int i = 0; i < n; ++i) sc[i] = codes[idx[i]];
	for (int i = 0; i < n; ++i) indices[i] = idx[i];
	if (codes_sorted) memcpy(codes_sorted, sc.data(), sizeof(uint32_t) * n);
	const uint32_t *C = sc.data();

	// build_tree_karras (lbvh-kernels.cu:186-256), one iteration per thread
	const int nn = 2 * n - 1;
	std::vector<uint32_t> par(nn, 0xffffffffu);
	for (int i = 0; i < nn; ++i) {
		rta_node &node = nodes[i];
		set_links(node, 0, 0);
		if (i >= n - 1) {
			set_links(node, 1u << 1, (uint32_t)(i - (n - 1)));
			continue;
		}
		int d = (cpl(i, i + 1, C, n) - cpl(i, i - 1, C, n)) >= 0 ? 1 : -1;
		int cpl_min = cpl(i, i - d, C, n);
		int l_max = 2;
		while (cpl(i, i + l_max * d, C, n) > cpl_min) l_max = l_max * 2;
		int l = 0;
		for (int t = l_max / 2; t >= 1; t = t / 2)
			if (cpl(i, i + (l + t) * d, C, n) > cpl_min) l = l + t;
		int j = i + l * d;
		int from = i, to = j;
		if (j < i) { from = j; to = i; }
		d = 1;
		int cpl_node = cpl(from, to, C, n);
		int s = 0;
		for (int t = l / 2 + l % 2; t >= 1; t = t / 2 + t % 2) {
			if (cpl(from, from + (s + t) * d, C, n) > cpl_node) s = s + t;
			if (t <= 1) break;
		}
		int mid = from + s * d + std::min(d, 0);
		int left = mid, right = mid + 1;
		if (from == mid) left = left + n - 1;
		if (to == mid + 1) right = right + n - 1;
		set_links(node, ((uint32_t)left << 1) | 1u, (uint32_t)right);
		par[left] = (uint32_t)i;
		par[right] = (uint32_t)i;
	}

	// build_bounding_boxes_karras2 (lbvh-kernels.cu:258-341), the leaf
	// threads run one after another.  A separate visit counter replaces the
	// reference's sum of codes (racy when a first arrival has code 0,
	// :288-291); the union is min(left, right) / max(left, right) with the
	// reference's (a<b)?a:b, fixing the sign of zero the reference leaves to
	// arrival order.
	auto rep = [&](uint32_t node) -> uint32_t {  // MORTON_CODE, :261
		return node >= (uint32_t)(n - 1) ? C[node - (n - 1)] : C[node];
	};
	std::vector<uint32_t> visit(n > 1 ? n - 1 : 1, 0);
	for (int k = 0; k < n; ++k) {
		uint32_t current = (uint32_t)(k + n - 1);
		rta_aabb box = boxes[idx[nodes[current].right_tris]];
		nodes[current].box = box;
		while (current != 0) {
			uint32_t last = current;
			uint32_t code = rep(current);
			current = par[current];
			if (visit[current]++ == 0) break;
			uint32_t L = nodes[current].type_left_elems >> 1, R = nodes[current].right_tris;
			rta_aabb box2 = (L == last) ? nodes[R].box : nodes[L].box;
			const rta_aabb &lb = (L == last) ? box : box2;
			const rta_aabb &rb = (L == last) ? box2 : box;
			rta_aabb u;
			u.min.x = MINF(lb.min.x, rb.min.x);
			u.min.y = MINF(lb.min.y, rb.min.y);
			u.min.z = MINF(lb.min.z, rb.min.z);
			u.max.x = MAXF(lb.max.x, rb.max.x);
			u.max.y = MAXF(lb.max.y, rb.max.y);
			u.max.z = MAXF(lb.max.z, rb.max.z);
			box = u;
			nodes[current].box = box;
			uint32_t old = rep(last == L ? R : L);
			if (code == old) {
				const rta_node &nl = nodes[L], &nr = nodes[R];
				if (!(nr.type_left_elems & 1u) && !(nl.type_left_elems & 1u)) {
					uint32_t tris = std::min(nr.right_tris, nl.right_tris);
					uint32_t length = (nr.type_left_elems >> 1) + (nl.type_left_elems >> 1);
					set_links(nodes[current], length << 1, tris);
				}
			}
		}
	}
	// invert_triangle_permutation_cuda (lbvh-kernels.cu:557-578)
	for (int i = 0; i < n; ++i) tris_out[i] = scratch[idx[i]];
	if (parents) memcpy(parents, par.data(), sizeof(uint32_t) * nn);
	return status;
}

// ---------------------------------------------------------- median ctor

// bbvh_constructor_using_median::build with spatial_median (the CPU
// plugin's default, bbvh/main.cpp:117; bbvh/bbvh.h:259-277).  nodes needs
// room for 2n-1 records.  Returns the node count; tris_out[n] receives the
// triangles in leaf order and order[n] their input indices.
int oracle_median_build(const rta_triangle *tris_in, int n, rta_node *nodes_out,
                        rta_triangle *tris_out, uint32_t *order) {
	if (n <= 0) return -1;
	MedianBuilder mb;
	mb.in = tris_in;
	mb.perm.resize(n);
	for (int i = 0; i < n; ++i) mb.perm[i] = (uint32_t)i;
	for (int ax = 0; ax < 3; ++ax) {
		mb.cog[ax].resize(n);
		for (int i = 0; i < n; ++i) mb.cog[ax][i] = cog_comp(tris_in[i], ax);
	}
	mb.nodes.reserve(2 * (size_t)n);
	mb.build_sm(0, (uint32_t)n);
	for (size_t i = 0; i < mb.nodes.size(); ++i) nodes_out[i] = mb.nodes[i];
	for (int i = 0; i < n; ++i) {
		tris_out[i] = tris_in[mb.out_order[i]];
		if (order) order[i] = mb.out_order[i];
	}
	return (int)mb.nodes.size();
}

// ------------------------------------------------------------- tracers

// bbvh_child_is_tracer::trace_rays / trace_ray, bbvh/bbvh.h:392-455.
// Rays are w*h with id y*w+x; rows are distributed with
// `schedule(dynamic, 32)` as in the reference.  Hits get the CPU tracer's
// values: closest.t starts at max_t and ref at 0 (traversal_state::reset,
// librta/raytrav.h:29-33); beta/gamma of a miss are left 0 here (the
// reference leaves them stale).  stats (optional, 2 words per ray) gets the
// number of inner nodes visited and of triangle tests.  first (optional)
// gets the first record the ray accepted (t = FLT_MAX if none).  Returns the
// wall time of the parallel loop in ms (bbvh.h:394,404).
double oracle_trace_cis(const rta_triangle *tris, const rta_node *nodes, const float *ro,
                        const float *rd, const float *rmax, int w, int h, rta_hit *hits,
                        uint32_t *stats, rta_hit *first, int threads) {
	if (threads > 0) omp_set_num_threads(threads);
	double t0 = now_ms();
#pragma omp parallel for schedule(dynamic, 32)
	for (int y = 0; y < h; ++y) {
		uint32_t stack[128];  // traversal_state::stack_size, raytrav.h:24
		for (int x = 0; x < w; ++x) {
			const size_t id = (size_t)y * w + x;
			const float *origin = ro + 3 * id, *dir = rd + 3 * id;
			const float max_t = rmax[id];
			rta_hit is_best;
			is_best.t = max_t;
			is_best.beta = 0;
			is_best.gamma = 0;
			is_best.ref = 0;
			uint32_t n_inner = 0, n_tris = 0;
			rta_hit first_rec;
			first_rec.t = FLT_MAX;
			first_rec.beta = 0;
			first_rec.gamma = 0;
			first_rec.ref = 0;
			bool accepted = false;
			int sp = 0;
			stack[0] = 0;
			while (sp >= 0) {
				uint32_t node = stack[sp--];
				const rta_node *curr = &nodes[node];
				if (curr->type_left_elems & 1u) {
					++n_inner;
					float dist_l = 0, dist_r = 0;
					uint32_t n_left = curr->type_left_elems >> 1;
					bool do_left = intersect_aabb(nodes[n_left].box, origin, dir, dist_l);
					uint32_t n_right = curr->right_tris;
					bool do_right = intersect_aabb(nodes[n_right].box, origin, dir, dist_r);
					do_left = do_left && dist_l < is_best.t && dist_l <= max_t;
					do_right = do_right && dist_r < is_best.t && dist_r <= max_t;
					if (do_left && do_right)
						if (dist_l <= dist_r) {
							stack[++sp] = n_right;
							stack[++sp] = n_left;
						} else {
							stack[++sp] = n_left;
							stack[++sp] = n_right;
						}
					else if (do_left)
						stack[++sp] = n_left;
					else if (do_right)
						stack[++sp] = n_right;
				} else {
					int elems = (int)(curr->type_left_elems >> 1);
					int offset = (int)curr->right_tris;
					for (int i = 0; i < elems; ++i) {
						++n_tris;
						float t, be, ga;
						if (intersect_tri_opt(tris[offset + i], origin, dir, t, be, ga)) {
							if (t < is_best.t && t <= max_t) {
								is_best.t = t;
								is_best.beta = be;
								is_best.gamma = ga;
								is_best.ref = (uint32_t)(offset + i);
								if (!accepted) first_rec = is_best;
								accepted = true;
							}
						}
					}
				}
			}
			hits[id] = is_best;
			if (first) first[id] = first_rec;
			if (stats) {
				stats[2 * id] = n_inner;
				stats[2 * id + 1] = n_tris;
			}
		}
	}
	return now_ms() - t0;
}

// bbvh_direct_is_tracer::trace_ray, bbvh/bbvh.h:330-374 (closest.t starts
// at FLT_MAX there, :349).
double oracle_trace_dis(const rta_triangle *tris, const rta_node *nodes, const float *ro,
                        const float *rd, const float *rmax, int w, int h, rta_hit *hits,
                        int threads) {
	if (threads > 0) omp_set_num_threads(threads);
	double t0 = now_ms();
#pragma omp parallel for schedule(dynamic, 32)
	for (int y = 0; y < h; ++y) {
		uint32_t stack[128];
		for (int x = 0; x < w; ++x) {
			const size_t id = (size_t)y * w + x;
			const float *origin = ro + 3 * id, *dir = rd + 3 * id;
			const float max_t = rmax[id];
			rta_hit best;
			best.t = FLT_MAX;
			best.beta = 0;
			best.gamma = 0;
			best.ref = 0;
			int sp = 0;
			stack[0] = 0;
			while (sp >= 0) {
				uint32_t node = stack[sp--];
				const rta_node *curr = &nodes[node];
				if (curr->type_left_elems & 1u) {
					float dist;
					if (intersect_aabb(curr->box, origin, dir, dist))
						if (dist < best.t && dist <= max_t) {
							stack[++sp] = curr->right_tris;
							stack[++sp] = curr->type_left_elems >> 1;
						}
				} else {
					int elems = (int)(curr->type_left_elems >> 1);
					int offset = (int)curr->right_tris;
					for (int i = 0; i < elems; ++i) {
						float t, be, ga;
						if (intersect_tri_opt(tris[offset + i], origin, dir, t, be, ga)) {
							if (t < best.t && t <= max_t) {
								best.t = t;
								best.beta = be;
								best.gamma = ga;
								best.ref = (uint32_t)(offset + i);
							}
						}
					}
				}
			}
			hits[id] = best;
		}
	}
	return now_ms() - t0;
}

// The GPU tracers' semantics over the same node array: k::trace_cis
// (bbvh-cuda/kernels.cu:777-855) and, with any_hit, k::trace_shadow_cis
// (:857-939).  closest starts as the incoming record with t = FLT_MAX
// (:793-794); at an inner node both child boxes are tested, a child is kept
// iff hit, dist < closest.t and dist <= max_t (:815-816); with both kept the
// nearer one (left on a tie) is entered and the other pushed (:818-827);
// with one kept it is entered; otherwise the stack is popped.  A leaf tests
// its triangles in order and keeps t < closest.t && t <= max_t (:838-846);
// the shadow form stops at the first kept triangle and then ends the ray
// (:918-931).  The reference's stack has 32 entries and no check; here the
// stack is larger and rays that would have needed more than 32 entries are
// counted in the return value; depth (optional) receives each ray's largest
// number of simultaneous stack entries.  hits: in/out.
int oracle_trace_kernel(const rta_triangle *tris, const rta_node *nodes, const float *ro,
                        const float *rd, const float *rmax, int n_rays, rta_hit *hits,
                        int any_hit, const float *bound, uint32_t *stats, uint32_t *depth,
                        int threads) {
	if (threads > 0) omp_set_num_threads(threads);
	int overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : overflow)
	for (int id = 0; id < n_rays; ++id) {
		uint32_t stack[256];
		const float *origin = ro + 3 * (size_t)id, *dir = rd + 3 * (size_t)id;
		const float t_max = rmax[id];
		// optional extra inclusive bound on child boxes (no reference
		// counterpart: the seeded traversal's prune, dist > bound)
		const float b = bound ? bound[id] : FLT_MAX;
		uint32_t n_inner = 0, n_tests = 0;
		rta_hit closest = hits[id];
		closest.t = FLT_MAX;
		int sp = -1, max_sp = -1;
		bool deep = false;
		const rta_node *curr = &nodes[0];
		while (true) {
			if (curr->type_left_elems & 1u) {
				++n_inner;
				const uint32_t l = curr->type_left_elems >> 1, r = curr->right_tris;
				float dist_l = FLT_MAX, dist_r = FLT_MAX;
				bool hit_l = intersect_aabb(nodes[l].box, origin, dir, dist_l);
				bool hit_r = intersect_aabb(nodes[r].box, origin, dir, dist_r);
				if (dist_l >= closest.t || dist_l > t_max) hit_l = false;
				if (dist_r >= closest.t || dist_r > t_max) hit_r = false;
				if (dist_l > b) hit_l = false;
				if (dist_r > b) hit_r = false;
				if (hit_l && hit_r) {
					if (sp + 1 >= 32) deep = true;
					if (sp + 1 >= 256) break;  // unreachable for trees of depth < 256
					if (dist_l <= dist_r) {
						stack[++sp] = r;
						curr = &nodes[l];
					} else {
						stack[++sp] = l;
						curr = &nodes[r];
					}
					if (sp > max_sp) max_sp = sp;
				} else if (hit_l) {
					curr = &nodes[l];
				} else if (hit_r) {
					curr = &nodes[r];
				} else if (sp >= 0) {
					curr = &nodes[stack[sp--]];
				} else {
					break;
				}
			} else {
				const uint32_t elems = curr->type_left_elems >> 1, offset = curr->right_tris;
				for (uint32_t i = 0; i < elems; ++i) {
					float t, be, ga;
					++n_tests;
					if (intersect_tri_opt(tris[offset + i], origin, dir, t, be, ga)) {
						if (t < closest.t && t <= t_max) {
							closest.t = t;
							closest.beta = be;
							closest.gamma = ga;
							closest.ref = offset + i;
							if (any_hit) break;
						}
					}
				}
				if (any_hit && closest.t != FLT_MAX) break;
				if (sp >= 0)
					curr = &nodes[stack[sp--]];
				else
					break;
			}
		}
		hits[id] = closest;
		if (stats) {
			stats[2 * (size_t)id] = n_inner;
			stats[2 * (size_t)id + 1] = n_tests;
		}
		if (depth) depth[id] = (uint32_t)(max_sp + 1);
		if (deep) ++overflow;
	}
	return overflow;
}

// bruteforce_tracer::trace_ray, brute-force/bruteforce.cpp:162-173: first
// minimum in index order, no max_t.
void oracle_trace_brute(const rta_triangle *tris, int n_tris, const float *ro, const float *rd,
                        int n_rays, rta_hit *hits) {
#pragma omp parallel for schedule(dynamic, 64)
	for (int r = 0; r < n_rays; ++r) {
		rta_hit best;
		best.t = FLT_MAX;
		best.beta = 0;
		best.gamma = 0;
		best.ref = 0;
		for (int i = 0; i < n_tris; ++i) {
			float t, be, ga;
			if (intersect_tri_opt(tris[i], ro + 3 * (size_t)r, rd + 3 * (size_t)r, t, be, ga))
				if (t < best.t) {
					best.t = t;
					best.beta = be;
					best.gamma = ga;
					best.ref = (uint32_t)i;
				}
		}
		hits[r] = best;
	}
}

// ------------------------------------------------------ diffuse bounce

static inline uint64_t mix64(uint64_t z) {
	z += 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

// The repo's incoherent ray set (no counterpart in the reference, whose
// bouncers never trace further bounces, librta/cuda.h:232-234): for each hit
// ray in ray order, origin = o + t*d + 1e-4*n with n the unit geometric
// normal ((b-a) x (c-a)) turned against d, direction = normalize(n + s)
// with s uniform on the sphere by Marsaglia's method, max_t = FLT_MAX.
// Returns the number of bounce rays.
int oracle_diffuse_bounce(const rta_hit *hits, const rta_triangle *tris, const float *ro,
                          const float *rd, int n_rays, uint64_t seed, float *out_o, float *out_d,
                          float *out_t, uint32_t *src_index) {
	int k = 0;
	for (int i = 0; i < n_rays; ++i) {
		const rta_hit &h = hits[i];
		if (h.t == FLT_MAX) continue;
		const rta_triangle &tr = tris[h.ref];
		const float e1x = tr.b.x - tr.a.x, e1y = tr.b.y - tr.a.y, e1z = tr.b.z - tr.a.z;
		const float e2x = tr.c.x - tr.a.x, e2y = tr.c.y - tr.a.y, e2z = tr.c.z - tr.a.z;
		float nx = e1y * e2z - e1z * e2y;
		float ny = e1z * e2x - e1x * e2z;
		float nz = e1x * e2y - e1y * e2x;
		const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
		nx = nx / nl;
		ny = ny / nl;
		nz = nz / nl;
		const float *o = ro + 3 * (size_t)i, *d = rd + 3 * (size_t)i;
		if (nx * d[0] + ny * d[1] + nz * d[2] > 0.0f) {
			nx = -nx;
			ny = -ny;
			nz = -nz;
		}
		const float px = o[0] + h.t * d[0], py = o[1] + h.t * d[1], pz = o[2] + h.t * d[2];
		out_o[3 * k + 0] = px + 1e-4f * nx;
		out_o[3 * k + 1] = py + 1e-4f * ny;
		out_o[3 * k + 2] = pz + 1e-4f * nz;
		const uint64_t base = mix64(seed ^ mix64((uint64_t)i));
		float x1 = 0, x2 = 0, s = 0;
		for (uint32_t it = 0; it < 64; ++it) {
			const uint64_t r = mix64(base + it);
			const float u1 = (float)(uint32_t)(r >> 40) * (1.0f / 16777216.0f);
			const float u2 = (float)(uint32_t)((r >> 16) & 0xffffffu) * (1.0f / 16777216.0f);
			const float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
			const float q = a * a + b * b;
			if (q < 1.0f) {
				x1 = a;
				x2 = b;
				s = q;
				break;
			}
		}
		const float sq = sqrtf(1.0f - s);
		const float sx = (2.0f * x1) * sq, sy = (2.0f * x2) * sq, sz = 1.0f - 2.0f * s;
		float vx = nx + sx, vy = ny + sy, vz = nz + sz;
		const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
		if (vl > 0.0f) {
			vx = vx / vl;
			vy = vy / vl;
			vz = vz / vl;
		} else {
			vx = nx;
			vy = ny;
			vz = nz;
		}
		out_d[3 * k + 0] = vx;
		out_d[3 * k + 1] = vy;
		out_d[3 * k + 2] = vz;
		out_t[k] = FLT_MAX;
		src_index[k] = (uint32_t)i;
		++k;
	}
	return k;
}

// Shadow rays towards a point light, the repo's own definition (as
// rta_shadow_rays on the GPU): origin as for the bounce, unit direction to
// the light, max_t = distance (0 with direction n for a light at the
// origin).  Returns the number of rays.
int oracle_shadow_rays(const rta_hit *hits, const rta_triangle *tris, const float *ro,
                       const float *rd, int n_rays, const float *light, float *out_o,
                       float *out_d, float *out_t, uint32_t *src_index) {
	int k = 0;
	for (int i = 0; i < n_rays; ++i) {
		const rta_hit &h = hits[i];
		if (h.t == FLT_MAX) continue;
		const rta_triangle &tr = tris[h.ref];
		const float e1x = tr.b.x - tr.a.x, e1y = tr.b.y - tr.a.y, e1z = tr.b.z - tr.a.z;
		const float e2x = tr.c.x - tr.a.x, e2y = tr.c.y - tr.a.y, e2z = tr.c.z - tr.a.z;
		float nx = e1y * e2z - e1z * e2y;
		float ny = e1z * e2x - e1x * e2z;
		float nz = e1x * e2y - e1y * e2x;
		const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
		nx = nx / nl;
		ny = ny / nl;
		nz = nz / nl;
		const float *o = ro + 3 * (size_t)i, *d = rd + 3 * (size_t)i;
		if (nx * d[0] + ny * d[1] + nz * d[2] > 0.0f) {
			nx = -nx;
			ny = -ny;
			nz = -nz;
		}
		const float px = o[0] + h.t * d[0], py = o[1] + h.t * d[1], pz = o[2] + h.t * d[2];
		const float sx = px + 1e-4f * nx, sy = py + 1e-4f * ny, sz = pz + 1e-4f * nz;
		out_o[3 * k + 0] = sx;
		out_o[3 * k + 1] = sy;
		out_o[3 * k + 2] = sz;
		const float vx = light[0] - sx, vy = light[1] - sy, vz = light[2] - sz;
		const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
		if (vl > 0.0f) {
			out_d[3 * k + 0] = vx / vl;
			out_d[3 * k + 1] = vy / vl;
			out_d[3 * k + 2] = vz / vl;
			out_t[k] = vl;
		} else {
			out_d[3 * k + 0] = nx;
			out_d[3 * k + 1] = ny;
			out_d[3 * k + 2] = nz;
			out_t[k] = 0.0f;
		}
		src_index[k] = (uint32_t)i;
		++k;
	}
	return k;
}

// Per-frame vertex jitter of the dynamic scene, the repo's own definition
// (as rta_jitter_triangles on the GPU): every vertex moves by sigma times an
// Irwin-Hall(4) variate per axis (four 16-bit uniforms of one hash keyed on
// the frame and vertex id, centred and scaled to unit variance).
void oracle_jitter_triangles(rta_triangle *out, const float *verts, const uint32_t *idx,
                             int n_tris, float sigma, uint64_t seed, uint32_t frame) {
	const uint64_t key = mix64(seed ^ mix64(0x6a09e667f3bcc909ull + frame));
	const float k = 1.0f / 65536.0f;
	for (int i = 0; i < n_tris; ++i) {
		float p[9];
		for (int c = 0; c < 3; ++c) {
			const uint32_t v = idx[3 * (size_t)i + c];
			for (int a = 0; a < 3; ++a) {
				const uint64_t h = mix64(key + 3ull * v + (uint64_t)a);
				const float u0 = (float)(uint32_t)(h & 0xffffu) * k;
				const float u1 = (float)(uint32_t)((h >> 16) & 0xffffu) * k;
				const float u2 = (float)(uint32_t)((h >> 32) & 0xffffu) * k;
				const float u3 = (float)(uint32_t)(h >> 48) * k;
				const float off = sigma * ((((u0 + u1) + (u2 + u3)) - 2.0f) * 1.7320508f);
				p[3 * c + a] = verts[3 * (size_t)v + a] + off;
			}
		}
		rta_triangle &t = out[i];
		memset(&t, 0, sizeof t);
		t.a = {p[0], p[1], p[2]};
		t.b = {p[3], p[4], p[5]};
		t.c = {p[6], p[7], p[8]};
		t.material_index = -1;
	}
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
