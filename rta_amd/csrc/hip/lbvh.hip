// Karras LBVH build for gfx950: the seven stages of the reference builder
// (bbvh-cuda/lbvh.h:153-211, bbvh-cuda/lbvh-kernels.cu) as hand-written HIP
// kernels.  Results are bit-identical to the reference's semantics; where
// the reference is racy the behaviour chosen here is deterministic and
// documented at the stage body (lbvh_body.hpp).  This file holds the
// single-mesh kernels and entry points; batch.hip runs the same bodies for
// many meshes per launch.
#include "lbvh_body.hpp"

#include <cstdio>

namespace rta_hip {

// One block per 256 triangles; its partial scene box goes to
// partials[blockIdx.x].
__global__ __launch_bounds__(kInitBlock) void init_triangles_kernel(
    rta_triangle *__restrict__ scratch, const rta_triangle *__restrict__ src,
    rta_aabb *__restrict__ boxes, rta_float3 *__restrict__ centers,
    rta_aabb *__restrict__ partials, int n) {
	init_triangles_block(scratch, src, boxes, centers, partials + blockIdx.x, n, blockIdx.x);
}

__global__ __launch_bounds__(kBoundsBlock) void voxel_bounds_kernel(
    rta_aabb *__restrict__ bounds, const rta_aabb *__restrict__ partials, int np) {
	voxel_bounds_block(bounds, partials, np);
}

__global__ __launch_bounds__(256) void morton_kernel(
    uint32_t *__restrict__ codes, uint32_t *__restrict__ indices,
    const rta_float3 *__restrict__ centers, const rta_aabb *__restrict__ bounds,
    uint32_t *__restrict__ status, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t code = morton_of(centers, bounds, i);
	codes[i] = code;
	indices[i] = (uint32_t)i;
	if (code == 0u && status) atomicOr(status, (uint32_t)RTA_LBVH_CODE_ZERO);
}

__global__ __launch_bounds__(kKarrasBlock) void karras_kernel(
    rta_node *__restrict__ nodes, uint32_t *__restrict__ parents,
    uint32_t *__restrict__ visit, uint2 *__restrict__ ranges,
    const uint32_t *__restrict__ codes, int n) {
	__shared__ uint32_t window[kKarrasWindow];
	const int base = blockIdx.x * kKarrasBlock;
	const CodeWindow w = stage_codes(window, codes, n, base);
	const int i = base + threadIdx.x;
	if (i >= 2 * n - 1) return;
	karras_node(nodes, parents, visit, ranges, w, n, i);
}

__global__ __launch_bounds__(kChunk) void bottom_up_chunk_kernel(
    rta_node *__restrict__ nodes, const rta_aabb *__restrict__ boxes,
    const uint32_t *__restrict__ codes, const uint32_t *__restrict__ indices,
    uint2 *__restrict__ ranges, uint4 *__restrict__ slots, uint32_t *__restrict__ counts,
    uint32_t *__restrict__ status, int n, PackOut pk) {
	bottom_up_chunk(nodes, boxes, codes, indices, ranges,
	                slots + (size_t)blockIdx.x * kSlots, counts + blockIdx.x, status, n,
	                (int)blockIdx.x, pk);
}

// One block merges the frontier slots of up to kMergeFan chunks of the
// previous level; its own chunk is `chunk` leaves wide.
__global__ __launch_bounds__(kMergeThreads) void bottom_up_merge_kernel(
    rta_node *__restrict__ nodes, uint2 *__restrict__ ranges,
    const uint4 *__restrict__ in_slots, const uint32_t *__restrict__ in_counts, int n_in,
    uint4 *__restrict__ out_slots, uint32_t *__restrict__ out_counts,
    uint32_t *__restrict__ status, long long chunk, int n, PackOut pk) {
	const int g0 = blockIdx.x * kMergeFan;
	const int ng = min(kMergeFan, n_in - g0);
	const uint32_t lo = (uint32_t)(blockIdx.x * chunk);
	const uint32_t hi = (uint32_t)min((long long)n, (long long)(blockIdx.x + 1) * chunk);
	bottom_up_merge(nodes, ranges, in_slots + (size_t)g0 * kSlots,
	                in_counts + g0, ng,
	                out_slots ? out_slots + (size_t)blockIdx.x * kSlots : nullptr,
	                out_counts ? out_counts + blockIdx.x : nullptr, status, lo, hi, n, pk);
}

struct BoxesWork {
	uint32_t *visit;     // per inner node (rta_lbvh_build_tree clears them)
	uint2 *ranges;       // per inner node
	uint4 *slots[2];     // frontier entries
	uint32_t *counts[2];
	uint32_t *status;
	size_t total;
};

static BoxesWork carve_boxes(void *base, int n) {
	BoxesWork w{};
	size_t off = 0;
	char *p = reinterpret_cast<char *>(base);
	auto take = [&](size_t bytes) {
		char *r = p ? p + off : nullptr;
		off += (bytes + 255) & ~(size_t)255;
		return r;
	};
	const size_t nb = (size_t)n;
	const size_t chunks = (nb + kChunk - 1) / kChunk;
	w.visit = (uint32_t *)take(nb * 4);
	w.ranges = (uint2 *)take(nb * 8);
	for (int i = 0; i < 2; ++i) {
		w.slots[i] = (uint4 *)take(chunks * kSlots * 16);
		w.counts[i] = (uint32_t *)take(chunks * 4);
	}
	w.status = (uint32_t *)take(4);
	w.total = off;
	return w;
}

__global__ __launch_bounds__(kPermuteBlock) void permute_kernel(
    rta_triangle *__restrict__ out, const rta_triangle *__restrict__ scratch,
    const uint32_t *__restrict__ indices, int n, float4 *__restrict__ tri4) {
	permute_block(out, scratch, indices, n, (int)blockIdx.x, tri4);
}

// --------------------------------------------------------- workspace

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct LbvhWorkspace {
	rta_triangle *scratch;
	rta_aabb *boxes;
	rta_float3 *centers;
	rta_aabb *partials;
	rta_aabb *bounds;
	uint32_t *codes;
	void *boxes_work;
	void *sort_work;
	size_t sort_bytes;
	size_t total;
};

static LbvhWorkspace carve(void *base, int n) {
	LbvhWorkspace w{};
	size_t off = 0;
	char *p = reinterpret_cast<char *>(base);
	auto take = [&](size_t bytes) {
		char *r = p ? p + off : nullptr;
		off += align256(bytes);
		return r;
	};
	const size_t nb = (size_t)n;
	const size_t nparts = (nb + kInitBlock - 1) / kInitBlock;
	w.scratch = (rta_triangle *)take(nb * sizeof(rta_triangle));
	w.boxes = (rta_aabb *)take(nb * sizeof(rta_aabb));
	w.centers = (rta_float3 *)take(nb * sizeof(rta_float3));
	w.partials = (rta_aabb *)take(nparts * sizeof(rta_aabb));
	w.bounds = (rta_aabb *)take(sizeof(rta_aabb));
	w.codes = (uint32_t *)take(nb * 4);
	w.boxes_work = take(carve_boxes(nullptr, n).total);
	size_t sb = 0;
	rta_lbvh_sort_work_bytes(n, &sb);
	w.sort_work = take(sb);
	w.sort_bytes = sb;
	w.total = off;
	return w;
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" {

int rta_lbvh_init_triangles(rta_triangle *scratch, const rta_triangle *src, rta_aabb *boxes,
                            rta_float3 *centers, rta_aabb *partials, int n, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_init_triangles: n must be > 0");
	if (!scratch || !src || !boxes || !centers || !partials)
		return invalid("rta_lbvh_init_triangles: NULL pointer");
	hipLaunchKernelGGL(init_triangles_kernel, dim3(grid_for(n, kInitBlock)), dim3(kInitBlock), 0,
	                   as_stream(s), scratch, src, boxes, centers, partials, n);
	RTA_LAUNCH_CHECK("init_triangles_kernel");
	return RTA_OK;
}

int rta_lbvh_voxel_bounds(rta_aabb *bounds, const rta_aabb *partials, int n, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_voxel_bounds: n must be > 0");
	if (!bounds || !partials) return invalid("rta_lbvh_voxel_bounds: NULL pointer");
	const int np = (int)grid_for(n, kInitBlock);
	hipLaunchKernelGGL(voxel_bounds_kernel, dim3(1), dim3(kBoundsBlock), 0, as_stream(s), bounds,
	                   partials, np);
	RTA_LAUNCH_CHECK("voxel_bounds_kernel");
	return RTA_OK;
}

int rta_lbvh_morton_codes(uint32_t *codes, uint32_t *indices, const rta_float3 *centers,
                          const rta_aabb *bounds, uint32_t *status, int n, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_morton_codes: n must be > 0");
	if (!codes || !indices || !centers || !bounds)
		return invalid("rta_lbvh_morton_codes: NULL pointer");
	hipLaunchKernelGGL(morton_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(s), codes,
	                   indices, centers, bounds, status, n);
	RTA_LAUNCH_CHECK("morton_kernel");
	return RTA_OK;
}

int rta_lbvh_boxes_work_bytes(int n, size_t *bytes) {
	if (n <= 0 || !bytes) return invalid("rta_lbvh_boxes_work_bytes: bad arguments");
	*bytes = carve_boxes(nullptr, n).total;
	return RTA_OK;
}

int rta_lbvh_build_tree(rta_node *nodes, uint32_t *parents, void *work, const uint32_t *codes,
                        int n, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_build_tree: n must be > 0");
	if (!nodes || !parents || !codes || !work) return invalid("rta_lbvh_build_tree: NULL pointer");
	BoxesWork w = carve_boxes(work, n);
	// the bottom-up stage ORs RTA_LBVH_FRONTIER_OVERFLOW into this word
	const int rc = rta_hip_fill_u32(w.status, 0u, 1, s);
	if (rc != RTA_OK) return rc;
	hipLaunchKernelGGL(karras_kernel, dim3(grid_for(2LL * n - 1, kKarrasBlock)), dim3(kKarrasBlock), 0,
	                   as_stream(s), nodes, parents, w.visit, w.ranges, codes, n);
	RTA_LAUNCH_CHECK("karras_kernel");
	return RTA_OK;
}

}  // extern "C"

// The bottom-up stage, which also emits the hierarchy (every node whole,
// and the inner nodes' ranges); a frontier overflow sets
// RTA_LBVH_FRONTIER_OVERFLOW in *status (the workspace's own word when
// status is null).
static int build_boxes(rta_node *nodes, void *work, const rta_aabb *boxes, const uint32_t *codes,
                       const uint32_t *indices, int n, uint32_t *status, rta_stream s,
                       PackOut pk = PackOut{}) {
	BoxesWork w = carve_boxes(work, n);
	if (!status) {
		status = w.status;
		const int rc = rta_hip_fill_u32(status, 0u, 1, s);
		if (rc != RTA_OK) return rc;
	}
	hipStream_t st = as_stream(s);
	int groups = (int)grid_for(n, kChunk);
	hipLaunchKernelGGL(bottom_up_chunk_kernel, dim3(groups), dim3(kChunk), 0, st, nodes, boxes,
	                   codes, indices, w.ranges, w.slots[0], w.counts[0], status, n, pk);
	RTA_LAUNCH_CHECK("bottom_up_chunk_kernel");
	long long chunk = kChunk;
	int in = 0;
	while (groups > 1) {
		const int blocks = (groups + kMergeFan - 1) / kMergeFan;
		chunk *= kMergeFan;
		hipLaunchKernelGGL(bottom_up_merge_kernel, dim3(blocks), dim3(kMergeThreads), 0, st, nodes,
		                   w.ranges, w.slots[in], w.counts[in], groups,
		                   blocks > 1 ? w.slots[1 - in] : nullptr,
		                   blocks > 1 ? w.counts[1 - in] : nullptr, status, chunk, n, pk);
		RTA_LAUNCH_CHECK("bottom_up_merge_kernel");
		groups = blocks;
		in = 1 - in;
	}
	return RTA_OK;
}

extern "C" {

int rta_lbvh_build_boxes(rta_node *nodes, void *work, const rta_aabb *boxes,
                         const uint32_t *codes, const uint32_t *indices,
                         const uint32_t *parents, int n, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_build_boxes: n must be > 0");
	(void)parents;   // the climb derives the hierarchy from the codes
	if (!nodes || !boxes || !codes || !indices || !work)
		return invalid("rta_lbvh_build_boxes: NULL pointer");
	return build_boxes(nodes, work, boxes, codes, indices, n, nullptr, s);
}

int rta_lbvh_permute(rta_triangle *out, const rta_triangle *scratch, const uint32_t *indices,
                     int n, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_permute: n must be > 0");
	if (!out || !scratch || !indices) return invalid("rta_lbvh_permute: NULL pointer");
	hipLaunchKernelGGL(permute_kernel, dim3(grid_for(n, kPermuteBlock)), dim3(kPermuteBlock), 0,
	                   as_stream(s), out, scratch, indices, n, (float4 *)nullptr);
	RTA_LAUNCH_CHECK("permute_kernel");
	return RTA_OK;
}

int rta_lbvh_workspace_bytes(int n, size_t *bytes) {
	if (n <= 0 || !bytes) return invalid("rta_lbvh_workspace_bytes: bad arguments");
	*bytes = carve(nullptr, n).total;
	return RTA_OK;
}

int rta_lbvh_copy_ranges(uint32_t *ranges, const void *workspace, int n, rta_stream s) {
	if (n <= 0 || !ranges || !workspace) return invalid("rta_lbvh_copy_ranges: bad arguments");
	if (n == 1) return RTA_OK;
	LbvhWorkspace w = carve(const_cast<void *>(workspace), n);
	BoxesWork b = carve_boxes(w.boxes_work, n);
	RTA_HIP_TRY(hipMemcpyAsync(ranges, b.ranges, (size_t)(n - 1) * 8, hipMemcpyDeviceToDevice,
	                           as_stream(s)));
	return RTA_OK;
}

}  // extern "C"

// The seven stages (lbvh_constructor::buildCUDA, bbvh-cuda/lbvh.h:153-211).
// With pk set, the permutation also writes tri4 and runs before the
// bottom-up pass, which writes the pair records (and clears the leaf-end
// marks that collapses make interior), so the traversal layout needs no
// pass of its own; the times then follow that order.
static int build_impl(const char *name, rta_triangle *tris, int n, rta_node *nodes,
                      uint32_t *indices, void *workspace, size_t workspace_bytes,
                      uint32_t *status, rta_lbvh_times *times, rta_stream s, PackOut pk) {
	const ScopedRange range(name);
	// n = 0 is rejected, as the reference cannot build it (bbvh-cuda/lbvh.h:47)
	if (n <= 0) return invalid((std::string(name) + ": n must be > 0").c_str());
	if (!tris || !nodes || !indices || !workspace)
		return invalid((std::string(name) + ": NULL pointer").c_str());
	LbvhWorkspace w = carve(workspace, n);
	if (workspace_bytes < w.total) {
		set_error(std::string(name) + ": workspace too small");
		return RTA_E_WORKSPACE;
	}
	hipEvent_t ev[8] = {};
	if (times) {
		for (int i = 0; i < 8; ++i) RTA_HIP_TRY(hipEventCreate(&ev[i]));
		RTA_HIP_TRY(hipEventRecord(ev[0], as_stream(s)));
	}
	int rc = RTA_OK;
	auto mark = [&](int k) {
		if (times && rc == RTA_OK) {
			hipError_t e = hipEventRecord(ev[k], as_stream(s));
			if (e != hipSuccess) rc = hip_status(e, "hipEventRecord");
		}
	};
	const bool packed = pk.pairs != nullptr;
	auto permute = [&]() {
		if (rc != RTA_OK) return;
		hipLaunchKernelGGL(permute_kernel, dim3(grid_for(n, kPermuteBlock)), dim3(kPermuteBlock),
		                   0, as_stream(s), tris, w.scratch, indices, n, pk.tri4);
		const hipError_t e = hipGetLastError();
		if (e != hipSuccess) rc = hip_status(e, "permute_kernel");
	};
	if (status) rc = rta_hip_fill_u32(status, 0u, 1, s);
	if (rc == RTA_OK && packed) rc = rta_hip_fill_u32(pk.flags, 0u, 1, s);
	if (rc == RTA_OK) rc = rta_lbvh_init_triangles(w.scratch, tris, w.boxes, w.centers, w.partials, n, s);
	mark(1);
	if (rc == RTA_OK) rc = rta_lbvh_voxel_bounds(w.bounds, w.partials, n, s);
	mark(2);
	if (rc == RTA_OK) rc = rta_lbvh_morton_codes(w.codes, indices, w.centers, w.bounds, status, n, s);
	mark(3);
	if (rc == RTA_OK) rc = rta_lbvh_sort(w.codes, indices, n, w.sort_work, w.sort_bytes, s);
	mark(4);
	// the hierarchy is emitted by the bottom-up stage: no pass of its own
	mark(5);
	if (packed) {
		permute();
		mark(6);
		if (rc == RTA_OK)
			rc = build_boxes(nodes, w.boxes_work, w.boxes, w.codes, indices, n, status, s, pk);
		mark(7);
	} else {
		if (rc == RTA_OK)
			rc = build_boxes(nodes, w.boxes_work, w.boxes, w.codes, indices, n, status, s);
		mark(6);
		permute();
		mark(7);
	}
	if (times) {
		if (rc == RTA_OK) {
			hipError_t e = hipEventSynchronize(ev[7]);
			if (e != hipSuccess) rc = hip_status(e, "hipEventSynchronize");
		}
		if (rc == RTA_OK) {
			float *slots[7] = {&times->init_triangles, &times->voxel_bounds, &times->morton,
			                   &times->sort, &times->tree,
			                   packed ? &times->permute : &times->boxes,
			                   packed ? &times->boxes : &times->permute};
			for (int k = 0; k < 7; ++k) (void)hipEventElapsedTime(slots[k], ev[k], ev[k + 1]);
			(void)hipEventElapsedTime(&times->total, ev[0], ev[7]);
		}
		for (int i = 0; i < 8; ++i) (void)hipEventDestroy(ev[i]);
	}
	return rc;
}

extern "C" {

int rta_lbvh_build(rta_triangle *tris, int n, rta_node *nodes, uint32_t *indices, void *workspace,
                   size_t workspace_bytes, uint32_t *status, rta_lbvh_times *times,
                   rta_stream s) {
	RTA_FAULT_POINT("rta_lbvh_build");
	return build_impl("rta_lbvh_build", tris, n, nodes, indices, workspace, workspace_bytes,
	                  status, times, s, PackOut{});
}

int rta_lbvh_build_packed(rta_triangle *tris, int n, rta_node *nodes, uint32_t *indices,
                          rta_node_pair *pairs, float *tri4, uint32_t *flags, void *workspace,
                          size_t workspace_bytes, uint32_t *status, rta_lbvh_times *times,
                          rta_stream s) {
	RTA_FAULT_POINT("rta_lbvh_build_packed");
	if (!pairs || !tri4 || !flags) return invalid("rta_lbvh_build_packed: NULL layout pointer");
	PackOut pk;
	pk.pairs = pairs;
	pk.tri4 = reinterpret_cast<float4 *>(tri4);
	pk.flags = flags;
	return build_impl("rta_lbvh_build_packed", tris, n, nodes, indices, workspace,
	                  workspace_bytes, status, times, s, pk);
}

}  // extern "C"
