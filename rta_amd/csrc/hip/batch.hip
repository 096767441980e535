// Batched LBVH build: many independent meshes built together, each stage
// one launch over all meshes.  Every block of a batched launch works on one
// mesh through that mesh's base pointers and runs the single-mesh stage body
// (lbvh_body.hpp), so each mesh's nodes, indices and reordered triangles are
// byte-identical to a build of that mesh alone.
//
// Sort: four 8-bit passes over the codes carrying (mesh << 24 | local index),
// then one pass over that payload's mesh byte carrying the codes.  Both are
// stable, so the result is every mesh's own stable sort by code, meshes in
// order (hence at most 256 meshes and 2^24 triangles per mesh in a batch).
#include "internal.hpp"
#include "lbvh_body.hpp"

#include <algorithm>
#include <new>

namespace rta_hip {

constexpr int kMaxBatchMeshes = 256;
constexpr uint32_t kMaxBatchTris = 1u << 24;

__global__ __launch_bounds__(kInitBlock) void batch_init_kernel(
    const BatchItem *__restrict__ items, const BatchMesh *__restrict__ meshes,
    rta_triangle *__restrict__ scratch, const rta_triangle *__restrict__ src,
    rta_aabb *__restrict__ boxes, rta_float3 *__restrict__ centers,
    rta_aabb *__restrict__ partials) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	init_triangles_block(scratch + m.tri_off, src + m.tri_off, boxes + m.tri_off,
	                     centers + m.tri_off, partials + m.part_off + it.block, (int)m.n,
	                     (int)it.block);
}

__global__ __launch_bounds__(kBoundsBlock) void batch_bounds_kernel(
    const BatchMesh *__restrict__ meshes, rta_aabb *__restrict__ bounds,
    const rta_aabb *__restrict__ partials) {
	const BatchMesh m = meshes[blockIdx.x];
	voxel_bounds_block(bounds + blockIdx.x, partials + m.part_off,
	                   (int)((m.n + kInitBlock - 1) / kInitBlock));
}

__global__ __launch_bounds__(256) void batch_morton_kernel(
    const BatchItem *__restrict__ items, const BatchMesh *__restrict__ meshes,
    uint32_t *__restrict__ codes, uint32_t *__restrict__ payload,
    const rta_float3 *__restrict__ centers, const rta_aabb *__restrict__ bounds,
    uint32_t *__restrict__ status) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	const int i = (int)it.block * 256 + (int)threadIdx.x;
	if (i >= (int)m.n) return;
	const uint32_t code = morton_of(centers + m.tri_off, bounds + it.mesh, i);
	codes[m.tri_off + i] = code;
	payload[m.tri_off + i] = (it.mesh << 24) | (uint32_t)i;
	if (code == 0u && status) atomicOr(status, (uint32_t)RTA_LBVH_CODE_ZERO);
}

__global__ __launch_bounds__(256) void batch_local_index_kernel(uint32_t *__restrict__ idx,
                                                                long long n) {
	const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
	if (i < n) idx[i] &= 0xffffffu;
}

// The traversal layout of one mesh of the batch (pairs, tri4 and flags
// are the batch's arrays; none when pairs is null).
__device__ inline PackOut mesh_pack(const BatchMesh &m, uint32_t mesh, rta_node_pair *pairs,
                                    float4 *tri4, uint32_t *flags) {
	PackOut pk;
	if (pairs) {
		pk.pairs = pairs + m.node_off;
		pk.tri4 = tri4 + 3 * (size_t)m.tri_off;
		pk.flags = flags + mesh;
	}
	return pk;
}

__global__ __launch_bounds__(kChunk) void batch_chunk_kernel(
    const BatchItem *__restrict__ items, const BatchMesh *__restrict__ meshes,
    rta_node *__restrict__ nodes, const rta_aabb *__restrict__ boxes,
    const uint32_t *__restrict__ codes, const uint32_t *__restrict__ indices,
    uint2 *__restrict__ ranges, uint4 *__restrict__ slots, uint32_t *__restrict__ counts,
    uint32_t *__restrict__ status, rta_node_pair *pairs, float4 *tri4, uint32_t *flags) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	bottom_up_chunk(nodes + m.node_off, boxes + m.tri_off, codes + m.tri_off, indices + m.tri_off,
	                ranges + m.tri_off,
	                slots + (size_t)(m.chunk_off + it.block) * kSlots, counts + m.chunk_off + it.block,
	                status, (int)m.n, (int)it.block, mesh_pack(m, it.mesh, pairs, tri4, flags));
}

// Merge level `level` (>= 1): level-`level` groups of a mesh cover
// kChunk * kMergeFan^level leaves.
__global__ __launch_bounds__(kMergeThreads) void batch_merge_kernel(
    const BatchItem *__restrict__ items, const BatchMesh *__restrict__ meshes,
    rta_node *__restrict__ nodes, uint2 *__restrict__ ranges,
    const uint4 *__restrict__ in_slots, const uint32_t *__restrict__ in_counts,
    uint4 *__restrict__ out_slots, uint32_t *__restrict__ out_counts,
    uint32_t *__restrict__ status, long long width, rta_node_pair *pairs, float4 *tri4,
    uint32_t *flags) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	const uint32_t g0 = it.block * kMergeFan;
	const int ng = (int)min((uint32_t)kMergeFan, it.groups - g0);
	const uint32_t lo = (uint32_t)(it.block * width);
	const uint32_t hi = (uint32_t)min((long long)m.n, (long long)(it.block + 1) * width);
	const bool last = it.blocks == 1;
	bottom_up_merge(nodes + m.node_off, ranges + m.tri_off,
	                in_slots + (size_t)(m.chunk_off + g0) * kSlots, in_counts + m.chunk_off + g0, ng,
	                last ? nullptr : out_slots + (size_t)(m.chunk_off + it.block) * kSlots,
	                last ? nullptr : out_counts + m.chunk_off + it.block, status, lo, hi,
	                (int)m.n, mesh_pack(m, it.mesh, pairs, tri4, flags));
}

__global__ __launch_bounds__(kPermuteBlock) void batch_permute_kernel(
    const BatchItem *__restrict__ items, const BatchMesh *__restrict__ meshes,
    rta_triangle *__restrict__ out, const rta_triangle *__restrict__ scratch,
    const uint32_t *__restrict__ indices, float4 *__restrict__ tri4) {
	const BatchItem it = items[blockIdx.x];
	const BatchMesh m = meshes[it.mesh];
	permute_block(out + m.tri_off, scratch + m.tri_off, indices + m.tri_off, (int)m.n,
	              (int)it.block, tri4 ? tri4 + 3 * (size_t)m.tri_off : nullptr);
}

// each mesh's root as its parent's pair record would describe it
__global__ void batch_root_words_kernel(const BatchMesh *__restrict__ meshes, int n_meshes,
                                        const rta_node *__restrict__ nodes,
                                        uint32_t *__restrict__ root_words) {
	const int m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= n_meshes) return;
	const rta_node r = nodes[meshes[m].node_off];
	root_words[m] = link_inner(r.type_left_elems) ? 1u : (r.right_tris << 1);
}

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct BatchWork {
	rta_triangle *scratch;
	rta_aabb *boxes, *partials, *bounds;
	rta_float3 *centers;
	uint32_t *codes;
	uint2 *ranges;
	uint4 *slots[2];
	uint32_t *counts[2];
	uint32_t *bu_status;
	void *sort_work;
	size_t sort_bytes, total;
};

static BatchWork carve_batch(void *base, const BatchPlan &p) {
	BatchWork w{};
	size_t off = 0;
	char *b = reinterpret_cast<char *>(base);
	auto take = [&](size_t bytes) {
		char *r = b ? b + off : nullptr;
		off += align256(bytes);
		return r;
	};
	const size_t N = (size_t)p.total_tris;
	w.scratch = (rta_triangle *)take(N * sizeof(rta_triangle));
	w.boxes = (rta_aabb *)take(N * sizeof(rta_aabb));
	w.centers = (rta_float3 *)take(N * sizeof(rta_float3));
	w.partials = (rta_aabb *)take((size_t)p.total_parts * sizeof(rta_aabb));
	w.bounds = (rta_aabb *)take((size_t)p.n_meshes * sizeof(rta_aabb));
	w.codes = (uint32_t *)take(N * 4);
	w.ranges = (uint2 *)take(N * 8);
	for (int i = 0; i < 2; ++i) {
		w.slots[i] = (uint4 *)take((size_t)p.total_chunks * kSlots * 16);
		w.counts[i] = (uint32_t *)take((size_t)p.total_chunks * 4);
	}
	w.bu_status = (uint32_t *)take(4);
	w.sort_bytes = radix_work_bytes((int)N);
	w.sort_work = take(w.sort_bytes);
	w.total = off;
	return w;
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" {

int rta_batch_create(int n_meshes, const int *n_tris, const int *n_rays, rta_batch *out) {
	if (!out || !n_tris || n_meshes <= 0) return invalid("rta_batch_create: bad arguments");
	if (n_meshes > kMaxBatchMeshes) return invalid("rta_batch_create: at most 256 meshes");
	BatchPlan *p = new (std::nothrow) BatchPlan;
	if (!p) return invalid("rta_batch_create: out of host memory");
	p->n_meshes = n_meshes;
	p->meshes.resize(n_meshes);
	std::vector<BatchItem> items;
	long long tri = 0, node = 0, part = 0, chunk = 0, ray = 0;
	for (int m = 0; m < n_meshes; ++m) {
		const int n = n_tris[m];
		const int r = n_rays ? n_rays[m] : 0;
		if (n <= 0 || (uint32_t)n >= kMaxBatchTris || r < 0) {
			delete p;
			return invalid("rta_batch_create: each mesh needs 1 .. 2^24-1 triangles");
		}
		BatchMesh &bm = p->meshes[m];
		bm = BatchMesh{(uint32_t)tri, (uint32_t)n, (uint32_t)node, (uint32_t)part,
		               (uint32_t)chunk, (uint32_t)ray, (uint32_t)r, 0u};
		tri += n;
		node += 2LL * n - 1;
		part += (n + kInitBlock - 1) / kInitBlock;
		chunk += (n + kChunk - 1) / kChunk;
		ray += r;
		if (tri > 0x7fffffffLL || node > 0xffffffffLL || ray > 0x7fffffffLL) {
			delete p;
			return invalid("rta_batch_create: batch too large");
		}
	}
	p->total_tris = tri; p->total_nodes = node; p->total_parts = part;
	p->total_chunks = chunk; p->total_rays = ray;
	// work lists
	p->tri_blocks_off = items.size();
	for (int m = 0; m < n_meshes; ++m)
		for (uint32_t b = 0; b * 256u < p->meshes[m].n; ++b) items.push_back({(uint32_t)m, b, 0, 0});
	p->tri_blocks = items.size() - p->tri_blocks_off;
	p->node_blocks_off = items.size();
	for (int m = 0; m < n_meshes; ++m)
		for (uint32_t b = 0; (long long)b * kKarrasBlock < 2LL * p->meshes[m].n - 1; ++b)
			items.push_back({(uint32_t)m, b, 0, 0});
	p->node_blocks = items.size() - p->node_blocks_off;
	p->chunk_items_off = items.size();
	for (int m = 0; m < n_meshes; ++m)
		for (uint32_t c = 0; (long long)c * kChunk < p->meshes[m].n; ++c)
			items.push_back({(uint32_t)m, c, 0, 0});
	p->chunk_items = items.size() - p->chunk_items_off;
	// merge levels: a mesh with g > 1 groups at a level merges them in
	// ceil(g / kMergeFan) blocks
	std::vector<uint32_t> groups(n_meshes);
	for (int m = 0; m < n_meshes; ++m) groups[m] = (p->meshes[m].n + kChunk - 1) / kChunk;
	while (true) {
		const size_t off = items.size();
		for (int m = 0; m < n_meshes; ++m) {
			if (groups[m] <= 1) continue;
			const uint32_t blocks = (groups[m] + kMergeFan - 1) / kMergeFan;
			for (uint32_t b = 0; b < blocks; ++b) items.push_back({(uint32_t)m, b, groups[m], blocks});
			groups[m] = blocks;
		}
		if (items.size() == off) break;
		p->merge_off.push_back(off);
		p->merge_len.push_back(items.size() - off);
	}
	p->n_items = items.size();
	hipError_t e = hipMalloc(&p->d_meshes, sizeof(BatchMesh) * n_meshes);
	if (e == hipSuccess) e = hipMalloc(&p->d_items, sizeof(BatchItem) * std::max<size_t>(1, items.size()));
	if (e == hipSuccess) e = hipMalloc(&p->d_root_words, sizeof(uint32_t) * n_meshes);
	if (e == hipSuccess) e = hipMalloc(&p->d_ovf, sizeof(uint32_t) * (size_t)(p->total_rays + 1));
	if (e == hipSuccess) e = hipMemset(p->d_ovf, 0, sizeof(uint32_t));
	if (e == hipSuccess)
		e = hipMemcpy(p->d_meshes, p->meshes.data(), sizeof(BatchMesh) * n_meshes,
		              hipMemcpyHostToDevice);
	if (e == hipSuccess && !items.empty())
		e = hipMemcpy(p->d_items, items.data(), sizeof(BatchItem) * items.size(),
		              hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		(void)hipFree(p->d_meshes);
		(void)hipFree(p->d_items);
		(void)hipFree(p->d_root_words);
		(void)hipFree(p->d_ovf);
		delete p;
		return hip_status(e, "rta_batch_create");
	}
	*out = (rta_batch)p;
	return RTA_OK;
}

int rta_batch_destroy(rta_batch b) {
	if (!b) return RTA_OK;
	BatchPlan *p = (BatchPlan *)b;
	hipError_t e1 = hipFree(p->d_meshes), e2 = hipFree(p->d_items), e3 = hipFree(p->d_root_words);
	hipError_t e4 = hipFree(p->d_ovf);
	delete p;
	if (e1 != hipSuccess) return hip_status(e1, "hipFree");
	if (e2 != hipSuccess) return hip_status(e2, "hipFree");
	if (e3 != hipSuccess) return hip_status(e3, "hipFree");
	if (e4 != hipSuccess) return hip_status(e4, "hipFree");
	return RTA_OK;
}

int rta_batch_overflow_count(rta_batch b, uint32_t *count, rta_stream s) {
	if (!b || !count) return invalid("rta_batch_overflow_count: NULL argument");
	const BatchPlan *p = (const BatchPlan *)b;
	hipStream_t st = as_stream(s);
	RTA_HIP_TRY(hipMemcpyAsync(count, p->d_ovf, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
	RTA_HIP_TRY(hipStreamSynchronize(st));
	return RTA_OK;
}

int rta_batch_sizes(rta_batch b, long long *total_tris, long long *total_nodes,
                    long long *total_rays) {
	if (!b) return invalid("rta_batch_sizes: NULL batch");
	const BatchPlan *p = (const BatchPlan *)b;
	if (total_tris) *total_tris = p->total_tris;
	if (total_nodes) *total_nodes = p->total_nodes;
	if (total_rays) *total_rays = p->total_rays;
	return RTA_OK;
}

int rta_batch_offsets(rta_batch b, uint32_t *tri_off, uint32_t *node_off, uint32_t *ray_off) {
	if (!b) return invalid("rta_batch_offsets: NULL batch");
	const BatchPlan *p = (const BatchPlan *)b;
	for (int m = 0; m < p->n_meshes; ++m) {
		if (tri_off) tri_off[m] = p->meshes[m].tri_off;
		if (node_off) node_off[m] = p->meshes[m].node_off;
		if (ray_off) ray_off[m] = p->meshes[m].ray_off;
	}
	return RTA_OK;
}

int rta_batch_workspace_bytes(rta_batch b, size_t *bytes) {
	if (!b || !bytes) return invalid("rta_batch_workspace_bytes: bad arguments");
	*bytes = carve_batch(nullptr, *(const BatchPlan *)b).total;
	return RTA_OK;
}

}  // extern "C"

// The stages of rta_batch_build; with pairs set, the traversal layout too
// (the permutation before the bottom-up pass writes tri4, the bottom-up
// pass the pair records and leaf-end marks, then each mesh's root word).
static int batch_build(const char *name, rta_batch b, rta_triangle *tris, rta_node *nodes,
                       uint32_t *indices, rta_node_pair *pairs, float4 *tri4, uint32_t *flags,
                       void *workspace, size_t workspace_bytes, uint32_t *status, rta_stream s) {
	if (!b || !tris || !nodes || !indices || !workspace)
		return invalid((std::string(name) + ": NULL argument").c_str());
	BatchPlan &p = *(BatchPlan *)b;
	BatchWork w = carve_batch(workspace, p);
	if (workspace_bytes < w.total) {
		set_error(std::string(name) + ": workspace too small");
		return RTA_E_WORKSPACE;
	}
	hipStream_t st = as_stream(s);
	const BatchItem *items = p.d_items;
	const int N = (int)p.total_tris;
	// the bottom-up stages OR RTA_LBVH_FRONTIER_OVERFLOW into the caller's
	// word, or into the workspace's own when there is none
	uint32_t *bu_status = status ? status : w.bu_status;
	{
		int r = rta_hip_fill_u32(bu_status, 0u, 1, s);
		if (r == RTA_OK && pairs) r = rta_hip_fill_u32(flags, 0u, (size_t)p.n_meshes, s);
		if (r != RTA_OK) return r;
	}
	hipLaunchKernelGGL(batch_init_kernel, dim3((unsigned)p.tri_blocks), dim3(kInitBlock), 0, st,
	                   items + p.tri_blocks_off, p.d_meshes, w.scratch, tris, w.boxes, w.centers,
	                   w.partials);
	RTA_LAUNCH_CHECK("batch_init_kernel");
	hipLaunchKernelGGL(batch_bounds_kernel, dim3((unsigned)p.n_meshes), dim3(kBoundsBlock), 0, st,
	                   p.d_meshes, w.bounds, w.partials);
	RTA_LAUNCH_CHECK("batch_bounds_kernel");
	hipLaunchKernelGGL(batch_morton_kernel, dim3((unsigned)p.tri_blocks), dim3(256), 0, st,
	                   items + p.tri_blocks_off, p.d_meshes, w.codes, indices, w.centers, w.bounds,
	                   status);
	RTA_LAUNCH_CHECK("batch_morton_kernel");
	int rc = radix_sort_pairs(w.codes, indices, N, 0, 4, w.sort_work, w.sort_bytes, st);
	if (rc != RTA_OK) return rc;
	rc = radix_sort_pairs(indices, w.codes, N, 24, 1, w.sort_work, w.sort_bytes, st);
	if (rc != RTA_OK) return rc;
	hipLaunchKernelGGL(batch_local_index_kernel, dim3(grid_for(N, 256)), dim3(256), 0, st,
	                   indices, (long long)N);
	RTA_LAUNCH_CHECK("batch_local_index_kernel");
	auto permute = [&]() -> int {
		hipLaunchKernelGGL(batch_permute_kernel, dim3((unsigned)p.tri_blocks), dim3(256), 0, st,
		                   items + p.tri_blocks_off, p.d_meshes, tris, w.scratch, indices, tri4);
		RTA_LAUNCH_CHECK("batch_permute_kernel");
		return RTA_OK;
	};
	if (pairs) {
		rc = permute();
		if (rc != RTA_OK) return rc;
	}
	// the bottom-up pass emits the hierarchy too (lbvh_body.hpp)
	hipLaunchKernelGGL(batch_chunk_kernel, dim3((unsigned)p.chunk_items), dim3(kChunk), 0, st,
	                   items + p.chunk_items_off, p.d_meshes, nodes, w.boxes, w.codes, indices,
	                   w.ranges, w.slots[0], w.counts[0], bu_status, pairs, tri4, flags);
	RTA_LAUNCH_CHECK("batch_chunk_kernel");
	long long width = kChunk;
	for (size_t L = 0; L < p.merge_off.size(); ++L) {
		width *= kMergeFan;
		const int in = (int)(L & 1), outb = 1 - in;
		hipLaunchKernelGGL(batch_merge_kernel, dim3((unsigned)p.merge_len[L]), dim3(kMergeThreads), 0,
		                   st, items + p.merge_off[L], p.d_meshes, nodes,
		                   w.ranges, w.slots[in], w.counts[in], w.slots[outb], w.counts[outb],
		                   bu_status, width, pairs, tri4, flags);
		RTA_LAUNCH_CHECK("batch_merge_kernel");
	}
	if (pairs) {
		hipLaunchKernelGGL(batch_root_words_kernel, dim3(grid_for(p.n_meshes, 64)), dim3(64), 0, st,
		                   p.d_meshes, p.n_meshes, nodes, p.d_root_words);
		RTA_LAUNCH_CHECK("batch_root_words_kernel");
		return RTA_OK;
	}
	return permute();
}

extern "C" {

int rta_batch_build(rta_batch b, rta_triangle *tris, rta_node *nodes, uint32_t *indices,
                    void *workspace, size_t workspace_bytes, uint32_t *status, rta_stream s) {
	return batch_build("rta_batch_build", b, tris, nodes, indices, nullptr, nullptr, nullptr,
	                   workspace, workspace_bytes, status, s);
}

int rta_batch_build_packed(rta_batch b, rta_triangle *tris, rta_node *nodes, uint32_t *indices,
                           rta_node_pair *pairs, float *tri4, uint32_t *flags, void *workspace,
                           size_t workspace_bytes, uint32_t *status, rta_stream s) {
	if (!pairs || !tri4 || !flags) return invalid("rta_batch_build_packed: NULL layout pointer");
	return batch_build("rta_batch_build_packed", b, tris, nodes, indices, pairs,
	                   reinterpret_cast<float4 *>(tri4), flags, workspace, workspace_bytes, status,
	                   s);
}

}  // extern "C"
