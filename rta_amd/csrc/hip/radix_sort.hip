// Stable LSD radix sort of 30-bit Morton codes with a uint32 payload
// (the reference's thrust::sort_by_key, bbvh-cuda/lbvh-kernels.cu:537-555).
//
// Reduce-then-scan, four 8-bit digit passes, each of three kernels:
//   1. per-tile digit histograms (LDS counters)           hist[digit][tile]
//   2. per-digit exclusive scan over tiles (one block per digit), digit totals
//   3. per-tile stable ranking with wave ballots (match-any over the digit
//      bits, popcount of the lower-lane mask) and LDS prefix counts, then a
//      digit-grouped write of the tile (contiguous runs per digit).
// No kernel waits on another block, so there is no spin loop anywhere.
// A stable sort is unique, so the output equals std::stable_sort by code.
#include "internal.hpp"

namespace rta_hip {

constexpr int kBits = 8;
constexpr int kBins = 1 << kBits;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kPasses = 4;  // 32 bits: covers the 30-bit codes

static_assert(kBins == kThreads, "one digit per thread in the scan helpers");

__device__ inline uint64_t lanemask_lt() {
	const int lane = __lane_id();
	return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// exclusive scan of one value per thread over the block (256 threads)
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *wave_sums, uint32_t &total) {
	const int lane = __lane_id();
	const int wv = threadIdx.x / kWave;
	uint32_t x = v;
	for (int off = 1; off < kWave; off <<= 1) {
		uint32_t y = __shfl_up(x, off);
		if (lane >= off) x += y;
	}
	if (lane == kWave - 1) wave_sums[wv] = x;
	__syncthreads();
	uint32_t before = 0, all = 0;
	for (int w = 0; w < kWaves; ++w) {
		uint32_t s = wave_sums[w];
		if (w < wv) before += s;
		all += s;
	}
	total = all;
	__syncthreads();
	return before + x - v;
}

__global__ __launch_bounds__(kThreads) void rs_hist_kernel(const uint32_t *__restrict__ keys, int n,
                                                           int shift, uint32_t *__restrict__ hist,
                                                           int tiles) {
	__shared__ uint32_t h[kBins];
	h[threadIdx.x] = 0;
	__syncthreads();
	const long long base = (long long)blockIdx.x * kTile;
#pragma unroll 4
	for (int it = 0; it < kItems; ++it) {
		const long long idx = base + it * kThreads + threadIdx.x;
		if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (kBins - 1)], 1u);
	}
	__syncthreads();
	hist[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// one block per digit: exclusive scan of hist[digit][0..tiles) in place
__global__ __launch_bounds__(kThreads) void rs_scan_kernel(uint32_t *__restrict__ hist, int tiles,
                                                           uint32_t *__restrict__ totals) {
	__shared__ uint32_t wave_sums[kWaves];
	uint32_t *row = hist + (size_t)blockIdx.x * tiles;
	uint32_t carry = 0;
	for (int base = 0; base < tiles; base += kThreads) {
		const int idx = base + threadIdx.x;
		const uint32_t v = idx < tiles ? row[idx] : 0u;
		uint32_t chunk_total;
		const uint32_t ex = block_exclusive_scan(v, wave_sums, chunk_total);
		if (idx < tiles) row[idx] = carry + ex;
		carry += chunk_total;
	}
	if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Per tile: rank every key among the tile's keys of its digit in index
// order (stable), then write the tile's keys grouped by digit so that each
// digit's run is stored contiguously.
//  1. every wave ranks its 64 keys of item `it` with ballots (match-any over
//     the digit bits, popcount of the lower lanes) and records the per-digit
//     count of (it, wave) in LDS;
//  2. one thread per digit turns the counts over (it, wave) -- the tile's
//     index order -- into exclusive prefixes and the digit's tile total;
//  3. a key's position in the tile is its digit's start in the tile plus
//     its (it, wave) prefix plus its rank; keys and payloads go to LDS there,
//     and the block writes them out in tile order, key j of digit d to
//     digit_base[d] + (j - start of d in the tile).
__global__ __launch_bounds__(kThreads) void rs_scatter_kernel(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, int n, int shift,
    const uint32_t *__restrict__ hist, const uint32_t *__restrict__ totals, int tiles) {
	__shared__ union {
		uint16_t cnt[kItems][kWaves][kBins];  // counts, then their prefixes
		uint32_t kv[2][kTile];                // the tile's keys and payloads by digit
	} sm;
	__shared__ uint32_t gbase[kBins];        // where digit d of this tile starts in the output
	__shared__ uint32_t lbase[kBins];        // where digit d starts within the tile
	__shared__ uint32_t wave_sums[kWaves];

	const int t = threadIdx.x;
	const int wv = t / kWave;
	uint32_t *cnt_words = reinterpret_cast<uint32_t *>(&sm.cnt[0][0][0]);
	for (int i = t; i < kItems * kWaves * kBins / 2; i += kThreads) cnt_words[i] = 0u;
	{
		uint32_t dummy;
		const uint32_t digit_base = block_exclusive_scan(totals[t], wave_sums, dummy);
		gbase[t] = digit_base + hist[(size_t)t * tiles + blockIdx.x];
	}
	__syncthreads();

	const uint64_t lt = lanemask_lt();
	const long long tile_base = (long long)blockIdx.x * kTile;
	const int count = (int)min((long long)kTile, (long long)n - tile_base);
	uint32_t key[kItems], val[kItems], pos[kItems];
#pragma unroll
	for (int it = 0; it < kItems; ++it) {
		const int j = it * kThreads + t;
		const bool valid = j < count;
		key[it] = valid ? keys_in[tile_base + j] : 0u;
		val[it] = valid ? vals_in[tile_base + j] : 0u;
		const uint32_t d = (key[it] >> shift) & (kBins - 1);
		uint64_t match = __ballot(valid);
#pragma unroll
		for (int b = 0; b < kBits; ++b) {
			const bool bit = (d >> b) & 1u;
			const uint64_t m = __ballot(bit);
			match &= bit ? m : ~m;
		}
		pos[it] = (uint32_t)__popcll(match & lt);
		if (valid && pos[it] == 0u) sm.cnt[it][wv][d] = (uint16_t)__popcll(match);
	}
	__syncthreads();
	uint32_t run = 0;
	for (int g = 0; g < kItems * kWaves; g += 16) {
		uint32_t c[16];
#pragma unroll
		for (int q = 0; q < 16; ++q) c[q] = sm.cnt[(g + q) / kWaves][(g + q) % kWaves][t];
#pragma unroll
		for (int q = 0; q < 16; ++q) {
			sm.cnt[(g + q) / kWaves][(g + q) % kWaves][t] = (uint16_t)run;
			run += c[q];
		}
	}
	{
		uint32_t dummy;
		lbase[t] = block_exclusive_scan(run, wave_sums, dummy);
	}
	__syncthreads();
#pragma unroll
	for (int it = 0; it < kItems; ++it) {
		const uint32_t d = (key[it] >> shift) & (kBins - 1);
		pos[it] += lbase[d] + sm.cnt[it][wv][d];
	}
	__syncthreads();  // the counts are dead; their LDS now holds the tile
#pragma unroll
	for (int it = 0; it < kItems; ++it) {
		if (it * kThreads + t < count) {
			sm.kv[0][pos[it]] = key[it];
			sm.kv[1][pos[it]] = val[it];
		}
	}
	__syncthreads();
	for (int j = t; j < count; j += kThreads) {
		const uint32_t k = sm.kv[0][j];
		const uint32_t d = (k >> shift) & (kBins - 1);
		const uint32_t o = gbase[d] + ((uint32_t)j - lbase[d]);
		if (o < (uint32_t)n) {  // always true for consistent histograms
			keys_out[o] = k;
			vals_out[o] = sm.kv[1][j];
		}
	}
}

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

size_t radix_work_bytes(int n) {
	const size_t tiles = ((size_t)n + kTile - 1) / kTile;
	return 2 * align256((size_t)n * 4) + align256(tiles * kBins * 4) + align256(kBins * 4);
}

int radix_sort_pairs(uint32_t *keys, uint32_t *vals, int n, int first_shift, int passes,
                     void *work, size_t work_bytes, hipStream_t st) {
	if (work_bytes < radix_work_bytes(n)) {
		set_error("radix sort: work buffer too small");
		return RTA_E_WORKSPACE;
	}
	const int tiles = (int)(((size_t)n + kTile - 1) / kTile);
	char *p = reinterpret_cast<char *>(work);
	uint32_t *alt_keys = reinterpret_cast<uint32_t *>(p);
	p += align256((size_t)n * 4);
	uint32_t *alt_vals = reinterpret_cast<uint32_t *>(p);
	p += align256((size_t)n * 4);
	uint32_t *hist = reinterpret_cast<uint32_t *>(p);
	p += align256((size_t)tiles * kBins * 4);
	uint32_t *totals = reinterpret_cast<uint32_t *>(p);

	uint32_t *kin = keys, *vin = vals, *kout = alt_keys, *vout = alt_vals;
	for (int pass = 0; pass < passes; ++pass) {
		const int shift = first_shift + pass * kBits;
		hipLaunchKernelGGL(rs_hist_kernel, dim3(tiles), dim3(kThreads), 0, st, kin, n, shift, hist,
		                   tiles);
		RTA_LAUNCH_CHECK("rs_hist_kernel");
		hipLaunchKernelGGL(rs_scan_kernel, dim3(kBins), dim3(kThreads), 0, st, hist, tiles, totals);
		RTA_LAUNCH_CHECK("rs_scan_kernel");
		hipLaunchKernelGGL(rs_scatter_kernel, dim3(tiles), dim3(kThreads), 0, st, kin, vin, kout,
		                   vout, n, shift, hist, totals, tiles);
		RTA_LAUNCH_CHECK("rs_scatter_kernel");
		uint32_t *tk = kin; kin = kout; kout = tk;
		uint32_t *tv = vin; vin = vout; vout = tv;
	}
	if (kin != keys) {
		// an odd number of passes ends in the alternate buffers
		RTA_HIP_TRY(hipMemcpyAsync(keys, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
		RTA_HIP_TRY(hipMemcpyAsync(vals, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
	}
	return RTA_OK;
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" {

int rta_lbvh_sort_work_bytes(int n, size_t *bytes) {
	if (n <= 0 || !bytes) return invalid("rta_lbvh_sort_work_bytes: bad arguments");
	*bytes = radix_work_bytes(n);
	return RTA_OK;
}

int rta_lbvh_sort(uint32_t *codes, uint32_t *indices, int n, void *work, size_t work_bytes,
                  rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_sort: n must be > 0");
	if (!codes || !indices || !work) return invalid("rta_lbvh_sort: NULL pointer");
	size_t need = 0;
	rta_lbvh_sort_work_bytes(n, &need);
	if (work_bytes < need) {
		set_error("rta_lbvh_sort: work buffer too small");
		return RTA_E_WORKSPACE;
	}
	// an even number of passes leaves the result in codes/indices
	static_assert(kPasses % 2 == 0, "result must end in the caller's buffers");
	return radix_sort_pairs(codes, indices, n, 0, kPasses, work, work_bytes, as_stream(s));
}

}  // extern "C"
