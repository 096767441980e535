// Per-frame geometry of the dynamic scene: every vertex of an indexed mesh
// moves by a seeded Gaussian offset (sigma per axis) and the triangle
// records the LBVH builder reads are written from the moved vertices.
#include "common.hpp"

namespace rta_hip {

// Offset of vertex `vid` along `axis` in frame `frame`: sigma times an
// Irwin-Hall(4) variate (four 16-bit uniforms from one hash, shifted to mean
// 0 and scaled to unit variance).  Plain binary32 operations in a fixed
// order, so the host restatement gives the same bits.
__device__ inline float vertex_offset(uint64_t frame_key, uint32_t vid, int axis, float sigma) {
	const uint64_t h = mix64(frame_key + 3ull * vid + (uint64_t)axis);
	const float k = 1.0f / 65536.0f;
	const float u0 = (float)(uint32_t)(h & 0xffffu) * k;
	const float u1 = (float)(uint32_t)((h >> 16) & 0xffffu) * k;
	const float u2 = (float)(uint32_t)((h >> 32) & 0xffffu) * k;
	const float u3 = (float)(uint32_t)(h >> 48) * k;
	return sigma * ((((u0 + u1) + (u2 + u3)) - 2.0f) * 1.7320508f);
}

__device__ __host__ inline uint64_t frame_key_of(uint64_t seed, uint32_t frame) {
	return mix64(seed ^ mix64(0x6a09e667f3bcc909ull + frame));
}

// frame_word (device, may be NULL) overrides `frame`, so a captured graph
// can advance the frame number itself (rta_hip_add_u32).  The block's
// records go through LDS, so that each store instruction writes one
// contiguous run of the output.
constexpr int kJitterBlock = 256;
constexpr int kTriWords32 = sizeof(rta_triangle) / 4;   // 26
__global__ __launch_bounds__(kJitterBlock) void jitter_triangles_kernel(
    rta_triangle *__restrict__ out, const float *__restrict__ verts,
    const uint32_t *__restrict__ idx, int n_tris, float sigma, uint64_t seed, uint32_t frame,
    const uint32_t *__restrict__ frame_word) {
	__shared__ uint32_t tile[kJitterBlock * kTriWords32];
	const int first = blockIdx.x * kJitterBlock;
	const int i = first + threadIdx.x;
	if (i < n_tris) {
		const uint64_t frame_key = frame_key_of(seed, frame_word ? *frame_word : frame);
		float p[9];
		for (int c = 0; c < 3; ++c) {
			const uint32_t v = idx[3 * (size_t)i + c];
			for (int a = 0; a < 3; ++a)
				p[3 * c + a] = verts[3 * (size_t)v + a] + vertex_offset(frame_key, v, a, sigma);
		}
		// the 104-byte record as 26 words: a, b, c; zero normals and
		// texcoords; material -1; the trailing alignment word 0 as well
		uint32_t *w = tile + threadIdx.x * kTriWords32;
		for (int k = 0; k < 9; ++k) w[k] = __float_as_uint(p[k]);
		for (int k = 9; k < 24; ++k) w[k] = 0u;
		w[24] = 0xffffffffu;
		w[25] = 0u;
	}
	__syncthreads();
	// the block's records are one run of count * 13 8-byte words
	const int count = min(kJitterBlock, n_tris - first);
	uint2 *dst = reinterpret_cast<uint2 *>(out + first);
	for (int k = threadIdx.x; k < count * (kTriWords32 / 2); k += kJitterBlock)
		dst[k] = make_uint2(tile[2 * k], tile[2 * k + 1]);
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" int rta_jitter_triangles(rta_triangle *out, const float *verts, const uint32_t *idx,
                                    int n_tris, float sigma, uint64_t seed, uint32_t frame,
                                    const uint32_t *frame_word, rta_stream s) {
	if (!out || !verts || !idx || n_tris <= 0 || !(sigma >= 0.0f))
		return invalid("rta_jitter_triangles: bad arguments");
	hipLaunchKernelGGL(jitter_triangles_kernel, dim3(grid_for(n_tris, kJitterBlock)),
	                   dim3(kJitterBlock), 0,
	                   as_stream(s), out, verts, idx, n_tris, sigma, seed, frame, frame_word);
	RTA_LAUNCH_CHECK("jitter_triangles_kernel");
	return RTA_OK;
}
