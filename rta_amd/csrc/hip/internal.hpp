// Internal interfaces shared between the translation units of librta_hip
// (not part of the C ABI).
#pragma once

#include "common.hpp"

#include <vector>

namespace rta_hip {

// Stable LSD radix sort of (keys, vals) by `passes` 8-bit digits starting at
// bit `first_shift` of the key; the result is left in keys/vals.  work:
// radix_work_bytes(n).
size_t radix_work_bytes(int n);
int radix_sort_pairs(uint32_t *keys, uint32_t *vals, int n, int first_shift, int passes,
                     void *work, size_t work_bytes, hipStream_t st);

// One mesh of a batch as the device sees it.
struct BatchMesh {
	uint32_t tri_off;    // first triangle in the concatenated triangle array
	uint32_t n;          // triangles
	uint32_t node_off;   // first node in the concatenated node array (2n-1 each)
	uint32_t part_off;   // first partial box (one per 256 triangles)
	uint32_t chunk_off;  // first level-1 bottom-up chunk (one per 1024 leaves)
	uint32_t ray_off;    // first ray in the concatenated ray arrays
	uint32_t n_rays;
	uint32_t pad;
};

// Work item of a batched launch: block `block` (in the stage's own units)
// of mesh `mesh`; merge levels also carry the mesh's group count at that
// level and its number of blocks.
struct BatchItem {
	uint32_t mesh, block, groups, blocks;
};

// The plan of a batch: mesh table and per-stage work lists on the device,
// their sizes on the host.
struct BatchPlan {
	int n_meshes = 0;
	long long total_tris = 0, total_nodes = 0, total_rays = 0;
	long long total_parts = 0, total_chunks = 0;
	std::vector<BatchMesh> meshes;            // host copy
	BatchMesh *d_meshes = nullptr;
	BatchItem *d_items = nullptr;             // all work lists, back to back
	size_t n_items = 0;
	// offsets / lengths of the work lists in d_items
	size_t tri_blocks_off = 0, tri_blocks = 0;     // 256 triangles per block
	size_t node_blocks_off = 0, node_blocks = 0;   // 256 nodes per block
	size_t chunk_items_off = 0, chunk_items = 0;   // 1024 leaves per block
	std::vector<size_t> merge_off, merge_len;      // per merge level
	uint32_t *d_root_words = nullptr;              // per mesh, written by the pack
	uint32_t *d_ovf = nullptr;    // short-stack overflow list: count, then up to total_rays rays
};

}  // namespace rta_hip
