// Learned hit predictor: a small MLP from an encoded ray to scores over the
// tree's treelets (a fixed cut into at most 1024 subtrees) plus a log hit
// distance, with hand-written BF16 MFMA kernels for training (FP32
// accumulate, FP32 master weights, Adam) and a fused inference kernel.  The
// predictor has no counterpart in the reference; it only proposes candidate
// treelets whose hit distance seeds a ray's max_t, and the exact traversal
// decides every result.
//
// Network (widths padded to multiples of 16 for the MFMA tiles):
//   x[112]  = 6 normalised inputs (origin mapped to [-1,1] by the scene box,
//             unit direction) and sin/cos(2^f pi v) for f = 0..7 of each
//             (102 features), zero padded
//   h1..h3  = ReLU(W h + b), 64 wide, bf16
//   out[1040] = W4 h3 + b4: 1024 treelet logits, the log hit distance at
//             column 1024, zero padding
// Weights are stored out-major ([out][in], the input index contiguous), so
// both MFMA operands of the training forward pass are one 8-byte load per
// lane.  This file is built without NaN handling (finite inputs only).
#include "common.hpp"

#include <cstring>
#include <new>
#include <vector>

namespace rta_hip {

namespace pred {

constexpr int kIn = 102, kInP = 112;
constexpr int kHid = 64;
constexpr int kClusters = 1024;
constexpr int kOut = kClusters + 1, kOutP = 1040;
constexpr int kTop = 4;

// parameter offsets (floats) in the packed master array
constexpr int oW1 = 0, oB1 = oW1 + kHid * kInP;
constexpr int oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid;
constexpr int oW4 = oB3 + kHid, oB4 = oW4 + kOutP * kHid;
constexpr int kParams = oB4 + kOutP;  // 83,152 with padding (81,537 used)

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (finite inputs)
__device__ __host__ inline uint16_t to_bf16(float f) {
	uint32_t u;
	memcpy(&u, &f, 4);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline v4s load4(const uint16_t *p) {
	const uint2 w = *reinterpret_cast<const uint2 *>(p);
	v4s r;
	r[0] = (short)(w.x & 0xffffu); r[1] = (short)(w.x >> 16);
	r[2] = (short)(w.y & 0xffffu); r[3] = (short)(w.y >> 16);
	return r;
}

// ---------------------------------------------------------- encoding

// sin and cos of 2^o pi v for o = o0..o0+3 (o0 = 0 or 4): sinpi/cospi at
// 2^o0 v (an exact scaling), then double-angle steps
// sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a); each step
// at most doubles the error, so the last octave is within ~1e-6 of the
// directly evaluated value.  e[2q] = sin, e[2q+1] = cos of octave o0+q.
// Training and inference share this, so both see identical inputs.
__device__ inline void encode_octaves(float v, int o0, float e[8]) {
	const float a = (o0 ? 16.0f : 1.0f) * v;
	float s = sinpif(a), c = cospif(a);
	for (int q = 0; q < 4; ++q) {
		e[2 * q] = s;
		e[2 * q + 1] = c;
		const float s2 = 2.0f * (s * c);
		c = (c - s) * (c + s);
		s = s2;
	}
}

__global__ __launch_bounds__(256) void encode_kernel(uint16_t *__restrict__ x,
                                                     const float *__restrict__ ro,
                                                     const float *__restrict__ rd, int n,
                                                     float3 lo, float3 scale) {
	const int r = blockIdx.x * 256 + threadIdx.x;
	if (r >= n) return;
	float v[6];
	v[0] = (ro[3 * (size_t)r + 0] - lo.x) * scale.x - 1.0f;
	v[1] = (ro[3 * (size_t)r + 1] - lo.y) * scale.y - 1.0f;
	v[2] = (ro[3 * (size_t)r + 2] - lo.z) * scale.z - 1.0f;
	v[3] = rd[3 * (size_t)r + 0];
	v[4] = rd[3 * (size_t)r + 1];
	v[5] = rd[3 * (size_t)r + 2];
	uint16_t *row = x + (size_t)r * kInP;
	for (int j = 0; j < 6; ++j) row[j] = to_bf16(v[j]);
	for (int j = 0; j < 6; ++j)
		for (int h = 0; h < 2; ++h) {
			float e[8];
			encode_octaves(v[j], 4 * h, e);
			for (int q = 0; q < 8; ++q) row[6 + 16 * j + 8 * h + q] = to_bf16(e[q]);
		}
	for (int j = kIn; j < kInP; ++j) row[j] = 0;
}

// ------------------------------------------------------------ GEMMs

// One wave, one 16x16 tile of Y = X W^T: X [m][k] row-major (ldx), W [nn][k]
// row-major (ldw); rows >= m read as zero.  Lane l holds the tile elements
// (4*(l/16) + i, l%16), i = 0..3.
__device__ inline v4f tile_xwt(const uint16_t *X, int ldx, const uint16_t *W, int ldw, int K,
                               int r0, int c0, int m) {
	const int lane = __lane_id();
	const int row = r0 + (lane & 15), col = c0 + (lane & 15), kq = 4 * (lane >> 4);
	v4f acc = {0.f, 0.f, 0.f, 0.f};
	for (int k0 = 0; k0 < K; k0 += 16) {
		v4s a = {0, 0, 0, 0};
		if (row < m) a = load4(X + (size_t)row * ldx + k0 + kq);
		const v4s b = load4(W + (size_t)col * ldw + k0 + kq);
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
	}
	return acc;
}

// Y = act(X W^T + b): one tile per wave, 4 waves per block.  OUT_F32: the
// last layer (f32, no activation); otherwise ReLU into bf16.
template <bool OUT_F32>
__global__ __launch_bounds__(256) void fwd_kernel(void *__restrict__ y, int ldy,
                                                  const uint16_t *__restrict__ x, int ldx,
                                                  const uint16_t *__restrict__ w,
                                                  const float *__restrict__ b, int K, int N,
                                                  int m) {
	const int tiles_n = N / 16;
	const int tile = blockIdx.x * 4 + threadIdx.x / kWave;
	const int tiles = ((m + 15) / 16) * tiles_n;
	if (tile >= tiles) return;
	const int r0 = (tile / tiles_n) * 16, c0 = (tile % tiles_n) * 16;
	const v4f acc = tile_xwt(x, ldx, w, K, K, r0, c0, m);
	const int lane = __lane_id();
	const int col = c0 + (lane & 15);
	for (int i = 0; i < 4; ++i) {
		const int row = r0 + 4 * (lane >> 4) + i;
		if (row >= m) continue;
		const float v = acc[i] + b[col];
		if (OUT_F32) reinterpret_cast<float *>(y)[(size_t)row * ldy + col] = v;
		else reinterpret_cast<uint16_t *>(y)[(size_t)row * ldy + col] = to_bf16(v > 0.f ? v : 0.f);
	}
}

// dX = (dY W) * [x_prev > 0]: dY [m][N] (lddy), W [N][K] row-major; tile of
// dX [m][K].  The B operand B[n][k] = W[n][k] is read with stride K.
__global__ __launch_bounds__(256) void bwd_data_kernel(uint16_t *__restrict__ dx,
                                                       const uint16_t *__restrict__ dy, int lddy,
                                                       const uint16_t *__restrict__ w, int K,
                                                       int N, const uint16_t *__restrict__ xprev,
                                                       int m) {
	const int tiles_k = K / 16;
	const int tile = blockIdx.x * 4 + threadIdx.x / kWave;
	const int tiles = ((m + 15) / 16) * tiles_k;
	if (tile >= tiles) return;
	const int r0 = (tile / tiles_k) * 16, c0 = (tile % tiles_k) * 16;
	const int lane = __lane_id();
	const int row = r0 + (lane & 15), col = c0 + (lane & 15), kq = 4 * (lane >> 4);
	v4f acc = {0.f, 0.f, 0.f, 0.f};
	for (int n0 = 0; n0 < N; n0 += 16) {
		v4s a = {0, 0, 0, 0};
		if (row < m) a = load4(dy + (size_t)row * lddy + n0 + kq);
		v4s bb;
		for (int i = 0; i < 4; ++i) bb[i] = (short)w[(size_t)(n0 + kq + i) * K + col];
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
	}
	for (int i = 0; i < 4; ++i) {
		const int r = r0 + 4 * (lane >> 4) + i;
		if (r >= m) continue;
		const size_t at = (size_t)r * K + col;
		dx[at] = xprev[at] != 0 ? to_bf16(acc[i]) : (uint16_t)0;
	}
}

// dW += dY^T X over a chunk of kRowChunk rows (fp32 atomics): dY [m][N]
// (lddy), X [m][K] (ldx); tile of dW [N][K].  Both operands are read with a
// row stride.
constexpr int kRowChunk = 256;
__global__ __launch_bounds__(256) void bwd_weight_kernel(float *__restrict__ dw,
                                                         const uint16_t *__restrict__ dy, int lddy,
                                                         const uint16_t *__restrict__ x, int ldx,
                                                         int K, int N, int m) {
	const int tiles_k = K / 16, tiles_n = N / 16;
	const int chunks = (m + kRowChunk - 1) / kRowChunk;
	const int tile = blockIdx.x * 4 + threadIdx.x / kWave;
	if (tile >= tiles_n * tiles_k * chunks) return;
	const int chunk = tile / (tiles_n * tiles_k);
	const int t2 = tile % (tiles_n * tiles_k);
	const int n0 = (t2 / tiles_k) * 16, k0 = (t2 % tiles_k) * 16;
	const int lane = __lane_id();
	const int nn = n0 + (lane & 15), kk = k0 + (lane & 15), rq = 4 * (lane >> 4);
	v4f acc = {0.f, 0.f, 0.f, 0.f};
	const int rend = min(m, (chunk + 1) * kRowChunk);
	for (int rb = chunk * kRowChunk; rb < rend; rb += 16) {
		v4s a, b;
		for (int i = 0; i < 4; ++i) {
			const int r = rb + rq + i;
			a[i] = r < m ? (short)dy[(size_t)r * lddy + nn] : (short)0;
			b[i] = r < m ? (short)x[(size_t)r * ldx + kk] : (short)0;
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
	}
	for (int i = 0; i < 4; ++i)
		atomicAdd(&dw[(size_t)(n0 + rq + i) * K + kk], acc[i]);
}

// db += column sums of dY: a block takes 64 columns of a kRowChunk-row
// chunk, its four waves every fourth row each (one coalesced 128-B row
// segment per wave and step), reduced in LDS; one atomic per column and
// block.
__global__ __launch_bounds__(256) void bwd_bias_kernel(float *__restrict__ db,
                                                       const uint16_t *__restrict__ dy, int lddy,
                                                       int N, int m) {
	__shared__ float part[4][kWave];
	const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
	const int col = blockIdx.x * kWave + lane;
	const int r0 = blockIdx.y * kRowChunk;
	const int r1 = min(m, r0 + kRowChunk);
	float s = 0.f;
	if (col < N)
		for (int r = r0 + w; r < r1; r += 4) s += from_bf16(dy[(size_t)r * lddy + col]);
	part[w][lane] = s;
	__syncthreads();
	if (w == 0 && col < N)
		atomicAdd(&db[col], (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
}

// --------------------------------------------------------- loss, top-k

// One wave per row: softmax cross-entropy over the 1024 treelet logits with
// the label's treelet and a squared error on the log distance, both for
// rows with a hit (label >= 0), averaged over the m rows.  Writes dOut
// (bf16, zero on padding columns and on rows without a hit) and adds the
// loss to *loss.
__global__ __launch_bounds__(256) void loss_kernel(uint16_t *__restrict__ dout,
                                                   const float *__restrict__ out,
                                                   const int32_t *__restrict__ label,
                                                   const float *__restrict__ target, int m,
                                                   float *__restrict__ loss) {
	const int row = blockIdx.x * 4 + threadIdx.x / kWave;
	if (row >= m) return;
	const int lane = __lane_id();
	const float *o = out + (size_t)row * kOutP;
	uint16_t *g = dout + (size_t)row * kOutP;
	const int lab = label[row];
	if (lab < 0) {
		for (int c = lane; c < kOutP; c += kWave) g[c] = 0;
		return;
	}
	// the lane's 16 logits stay in registers: max, then exp, then gradients
	float v[kClusters / kWave];
	float mx = -FLT_MAX;
#pragma unroll
	for (int j = 0; j < kClusters / kWave; ++j) {
		v[j] = o[lane + kWave * j];
		mx = fmaxf(mx, v[j]);
	}
	for (int off = kWave / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
	float se = 0.f;
#pragma unroll
	for (int j = 0; j < kClusters / kWave; ++j) {
		v[j] = expf(v[j] - mx);
		se += v[j];
	}
	for (int off = kWave / 2; off > 0; off >>= 1) se += __shfl_xor(se, off);
	const float inv_m = 1.0f / (float)m;
	const float inv_se = 1.0f / se;   // one division per row (softmax within 1 ulp)
#pragma unroll
	for (int j = 0; j < kClusters / kWave; ++j) {
		const int c = lane + kWave * j;
		g[c] = to_bf16((v[j] * inv_se - (c == lab ? 1.f : 0.f)) * inv_m);
	}
	if (lane < kOutP - kClusters) {
		const int c = kClusters + lane;
		g[c] = to_bf16(c == kClusters ? 2.f * (o[c] - target[row]) * inv_m : 0.f);
	}
	if (lane == 0) {
		const float d = o[kClusters] - target[row];
		atomicAdd(loss, ((logf(se) + mx - o[lab]) + d * d) * inv_m);
	}
}

// One wave per row: the kTop highest logits (ties: lower treelet first) and
// the predicted log distance.
__global__ __launch_bounds__(256) void topk_kernel(int32_t *__restrict__ top,
                                                   float *__restrict__ logdist,
                                                   const float *__restrict__ out, int m) {
	const int row = blockIdx.x * 4 + threadIdx.x / kWave;
	if (row >= m) return;
	const int lane = __lane_id();
	const float *o = out + (size_t)row * kOutP;
	float v[kClusters / kWave];
	for (int j = 0; j < kClusters / kWave; ++j) v[j] = o[lane + kWave * j];
	for (int t = 0; t < kTop; ++t) {
		float best = -FLT_MAX;
		int bi = 0x7fffffff;
		for (int j = 0; j < kClusters / kWave; ++j)
			if (v[j] > best) { best = v[j]; bi = lane + kWave * j; }
		for (int off = kWave / 2; off > 0; off >>= 1) {
			const float ob = __shfl_xor(best, off);
			const int oi = __shfl_xor(bi, off);
			if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
		}
		if (top && lane == 0) top[(size_t)row * kTop + t] = bi;
		if ((bi & (kWave - 1)) == lane) v[bi / kWave] = -FLT_MAX;
	}
	if (logdist && lane == 0) logdist[row] = o[kClusters];
}

// Adam on the packed parameters; refreshes the bf16 weight copies.
// step_word (optional): take the step from device memory (bias corrections
// computed here) instead of c1, c2, so a captured launch advances with it.
__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, float *__restrict__ mo,
                                                   float *__restrict__ ve,
                                                   const float *__restrict__ g,
                                                   uint16_t *__restrict__ wb, float lr, float c1,
                                                   float c2, const uint32_t *__restrict__ step_word) {
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kParams) return;
	if (step_word) {
		const float st = (float)*step_word;
		c1 = 1.0f / (1.0f - powf(0.9f, st));
		c2 = 1.0f / (1.0f - powf(0.999f, st));
	}
	const float gi = g[i];
	const float m1 = 0.9f * mo[i] + 0.1f * gi;
	const float v1 = 0.999f * ve[i] + 0.001f * gi * gi;
	mo[i] = m1;
	ve[i] = v1;
	const float pi = p[i] - lr * (m1 * c1) / (sqrtf(v1 * c2) + 1e-8f);
	p[i] = pi;
	wb[i] = to_bf16(pi);
}

// n_out rays (and their hit records) drawn from the first *count of a ray
// set: entry i takes ray hash(seed, frame, i) mod *count, frame = *frame_word
// (0 without one), so a captured launch draws a new sample every frame.  An
// empty set gives misses.
__global__ __launch_bounds__(256) void sample_kernel(
    const float *__restrict__ o, const float *__restrict__ d, const rta_hit *__restrict__ hits,
    const uint32_t *__restrict__ count, int n_out, const uint32_t *__restrict__ frame_word,
    uint64_t seed, float *__restrict__ o_out, float *__restrict__ d_out,
    rta_hit *__restrict__ hits_out) {
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n_out) return;
	const uint32_t c = *count;
	if (c == 0u) {
		for (int k = 0; k < 3; ++k) o_out[3 * (size_t)i + k] = d_out[3 * (size_t)i + k] = 0.0f;
		hits_out[i] = rta_hit{FLT_MAX, 1.0f, 2.0f, 0u};
		return;
	}
	uint64_t z = seed * 0x9e3779b97f4a7c15ull + ((uint64_t)(frame_word ? *frame_word : 0u) << 32) +
	             (uint64_t)i;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	z ^= z >> 31;
	const size_t j = (size_t)(z % c);
	for (int k = 0; k < 3; ++k) {
		o_out[3 * (size_t)i + k] = o[3 * j + k];
		d_out[3 * (size_t)i + k] = d[3 * j + k];
	}
	hits_out[i] = hits[j];
}

__global__ __launch_bounds__(256) void to_bf16_kernel(uint16_t *__restrict__ wb,
                                                      const float *__restrict__ p) {
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i < kParams) wb[i] = to_bf16(p[i]);
}

// ------------------------------------------------- fused inference
//
// One kernel from rays to top-4 treelets: activations stay in registers and
// the 1040 logits of a ray never reach memory.  The products are computed
// transposed, D = W H (rows = output features, columns = rays), with the
// 16x16x32 bf16 MFMA: lane l supplies weight row l%16 and ray l%16, and the
// 8 k-slots of lane group g = l/16.  Lane l then holds D[4g+i][l%16], which
// is already the next layer's B operand once the k-slots are assigned as
//   k-step s, group g, slot j<4  -> feature 32s + 4g + j       (tile 2s)
//                      slot j>=4 -> feature 32s + 16 + 4g + j-4 (tile 2s+1)
// so the weights are stored permuted to match (pack_infer_kernel), one 16-B
// record per (tile, k-step, lane).  The first layer's slots are the encoded
// inputs rotated by 6 (k' = feature - 6 for the 96 sin/cos features, then
// the 6 raw inputs at k' = 96..101): slot (s, g, j) is octave block 4s+g,
// which a lane computes by itself.
//
// Top-4 per ray: each logit becomes a key whose low 10 mantissa bits are
// replaced by 1023 - treelet (distinct keys; the order of the keys is the
// order of the logits truncated to 13 mantissa bits, ties to the lower
// treelet for positive logits), kept in a sorted register quadruple by one
// max and three med3 per key, and the four lanes of a ray merged by
// shuffles at the end.

typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2b __attribute__((ext_vector_type(2)));

constexpr int kInferWaves = 16;                       // 1024 threads per block
constexpr int kRec1 = 4 * 4 * 64;                     // layer 1: 4 tiles x 4 k-steps
constexpr int kRecH = 4 * 2 * 64;                     // layers 2, 3
constexpr int kRec4 = (kOutP / 16) * 2 * 64;          // layer 4: 65 tiles x 2 k-steps
constexpr int pW1 = 0, pW2 = pW1 + kRec1, pW3 = pW2 + kRecH, pW4 = pW3 + kRecH;
constexpr int kPackRecords = pW4 + kRec4;             // 16-B records (165,888 B)

__device__ inline unsigned bf16x2(float a, float b) {
	return __builtin_bit_cast(unsigned, __builtin_convertvector((v2f){a, b}, v2b));
}

__device__ inline v4f mfma32(v4u a, v4u b, v4f c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8b, a),
	                                               __builtin_bit_cast(v8b, b), c, 0, 0, 0);
}

// wp[e] for element e = 8 * record + slot of the permuted bf16 weights
__global__ __launch_bounds__(256) void pack_infer_kernel(uint16_t *__restrict__ wp,
                                                         const uint16_t *__restrict__ wb) {
	const int e = blockIdx.x * 256 + threadIdx.x;
	if (e >= kPackRecords * 8) return;
	const int rec = e >> 3, j = e & 7, lane = rec & 63, g = lane >> 4;
	uint16_t val = 0;
	if (rec < pW2) {
		const int r = rec - pW1, t = r / (4 * 64), s = (r / 64) % 4;
		const int kp = 32 * s + 8 * g + j;
		const int col = kp < 96 ? kp + 6 : (kp < 102 ? kp - 96 : -1);
		if (col >= 0) val = wb[oW1 + (16 * t + (lane & 15)) * kInP + col];
	} else {
		const int base = rec < pW3 ? pW2 : (rec < pW4 ? pW3 : pW4);
		const int off = rec < pW3 ? oW2 : (rec < pW4 ? oW3 : oW4);
		const int r = rec - base, t = r / (2 * 64), s = (r / 64) % 2;
		const int k = j < 4 ? 32 * s + 4 * g + j : 32 * s + 16 + 4 * g + (j - 4);
		val = wb[off + (16 * t + (lane & 15)) * kHid + k];
	}
	wp[e] = val;
}

// one 64-wide ReLU layer for two sets of 16 rays; in: KS k-steps of B
// operands per set, out: the next layer's 2 k-steps per set
template <int KS>
__device__ inline void hidden_layer(v4u out[2][2], const v4u in[2][KS], const v4u *__restrict__ wp,
                                    const float *__restrict__ bias, int lane) {
	unsigned d[2][4][2];
	#pragma unroll
	for (int t = 0; t < 4; ++t) {
		const float4 b = *reinterpret_cast<const float4 *>(bias + 16 * t + 4 * (lane >> 4));
		v4f a0 = {b.x, b.y, b.z, b.w}, a1 = a0;
		#pragma unroll
		for (int s = 0; s < KS; ++s) {
			const v4u w = wp[(t * KS + s) * 64 + lane];
			a0 = mfma32(w, in[0][s], a0);
			a1 = mfma32(w, in[1][s], a1);
		}
		d[0][t][0] = bf16x2(fmaxf(a0[0], 0.f), fmaxf(a0[1], 0.f));
		d[0][t][1] = bf16x2(fmaxf(a0[2], 0.f), fmaxf(a0[3], 0.f));
		d[1][t][0] = bf16x2(fmaxf(a1[0], 0.f), fmaxf(a1[1], 0.f));
		d[1][t][1] = bf16x2(fmaxf(a1[2], 0.f), fmaxf(a1[3], 0.f));
	}
	#pragma unroll
	for (int r = 0; r < 2; ++r)
		#pragma unroll
		for (int s = 0; s < 2; ++s)
			out[r][s] = (v4u){d[r][2 * s][0], d[r][2 * s][1], d[r][2 * s + 1][0], d[r][2 * s + 1][1]};
}

// sorted (t[0] > t[1] > t[2] > t[3]) insertion of a key
__device__ inline void insert4(float t[4], float x) {
	const float n0 = __builtin_amdgcn_fmed3f(t[0], x, __builtin_inff());
	const float n1 = __builtin_amdgcn_fmed3f(t[0], t[1], x);
	const float n2 = __builtin_amdgcn_fmed3f(t[1], t[2], x);
	const float n3 = __builtin_amdgcn_fmed3f(t[2], t[3], x);
	t[0] = n0; t[1] = n1; t[2] = n2; t[3] = n3;
}

__device__ inline int key_cluster(float k) { return 1023 - (int)(__float_as_uint(k) & 0x3ffu); }

// Persistent: each wave takes tiles of 32 rays (two sets of 16) from a
// static stride; layer 4's permuted weights (133,120 B) and biases sit in
// LDS, one block of 16 waves per CU.  LOGITS: also write every output (for
// testing; a separate instance keeps the stores out of the fast loop).  K:
// 4, or 1 (top-1 only: one max per key; entries 1..3 of a ray's top are
// written as -1).
template <bool LOGITS, int K>
__global__ __launch_bounds__(1024) void infer_kernel(const v4u *__restrict__ wp,
                                                     const float *__restrict__ params,
                                                     const float *__restrict__ ro,
                                                     const float *__restrict__ rd, int n,
                                                     float3 lo, float3 scale,
                                                     int32_t *__restrict__ top,
                                                     float *__restrict__ logdist,
                                                     float *__restrict__ logits) {
	__shared__ v4u w4s[kRec4];
	__shared__ v4f b4s[kOutP / 4];
	for (int i = threadIdx.x; i < kRec4; i += kInferWaves * kWave) w4s[i] = wp[pW4 + i];
	for (int i = threadIdx.x; i < kOutP / 4; i += kInferWaves * kWave)
		b4s[i] = reinterpret_cast<const v4f *>(params + oB4)[i];
	__syncthreads();
	const int lane = __lane_id(), g = lane >> 4, col = lane & 15;
	const int n_tiles = (n + 31) / 32;
	for (int tile = blockIdx.x * kInferWaves + (int)(threadIdx.x / kWave); tile < n_tiles;
	     tile += gridDim.x * kInferWaves) {
		// layer 1 operands: octave blocks 4s+g of the lane's ray, s = 0..2,
		// and the raw inputs (group 0) at s = 3
		v4u x[2][4];
		#pragma unroll
		for (int r = 0; r < 2; ++r) {
			const int ray = tile * 32 + r * 16 + col;
			float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
			if (ray < n) {
				v0 = (ro[3 * (size_t)ray + 0] - lo.x) * scale.x - 1.0f;
				v1 = (ro[3 * (size_t)ray + 1] - lo.y) * scale.y - 1.0f;
				v2 = (ro[3 * (size_t)ray + 2] - lo.z) * scale.z - 1.0f;
				v3 = rd[3 * (size_t)ray + 0];
				v4 = rd[3 * (size_t)ray + 1];
				v5 = rd[3 * (size_t)ray + 2];
			}
			// block 4s+g encodes input 2s + g/2, octaves 4*(g%2)..+3
			const float in[3] = {(g & 2) ? v1 : v0, (g & 2) ? v3 : v2, (g & 2) ? v5 : v4};
			#pragma unroll
			for (int s = 0; s < 3; ++s) {
				float e[8];
				encode_octaves(in[s], (g & 1) ? 4 : 0, e);
				x[r][s] = (v4u){bf16x2(e[0], e[1]), bf16x2(e[2], e[3]), bf16x2(e[4], e[5]),
				                bf16x2(e[6], e[7])};
			}
			x[r][3] = g == 0 ? (v4u){bf16x2(v0, v1), bf16x2(v2, v3), bf16x2(v4, v5), 0u}
			                 : (v4u){0u, 0u, 0u, 0u};
		}
		// the hidden layers' weights and biases are loaded per tile (from the
		// caches): the empty asm hides their loop invariance, which would
		// otherwise hoist all of them into registers and spill
		const v4u *wq = wp;
		const float *pq = params;
		asm volatile("" : "+s"(wq), "+s"(pq));
		v4u h1[2][2], h2[2][2], h3[2][2];
		hidden_layer<4>(h1, x, wq + pW1, pq + oB1, lane);
		hidden_layer<2>(h2, h1, wq + pW2, pq + oB2, lane);
		hidden_layer<2>(h3, h2, wq + pW3, pq + oB3, lane);
		// layer 4: 64 treelet tiles into the running top-4 of each lane,
		// software-pipelined: the products of tile t+1 are issued and the
		// operands of tile t+2 fetched before the keys of tile t are inserted;
		// tile 64 (the log distance and padding) comes out of the last round
		float tk[2][4];
		#pragma unroll
		for (int r = 0; r < 2; ++r)
			#pragma unroll
			for (int i = 0; i < 4; ++i) tk[r][i] = -__builtin_inff();
		const unsigned low0 = 1023u - 4u * (unsigned)g;
		v4u w0 = w4s[lane], w1 = w4s[64 + lane];
		v4f bias = b4s[g];
		v4f acc[2];
		acc[0] = mfma32(w1, h3[0][1], mfma32(w0, h3[0][0], bias));
		acc[1] = mfma32(w1, h3[1][1], mfma32(w0, h3[1][0], bias));
		w0 = w4s[2 * 64 + lane];
		w1 = w4s[3 * 64 + lane];
		bias = b4s[4 + g];
#pragma unroll 1
		for (int t = 0; t < kClusters / 16; ++t) {
			v4f nxt[2];
			nxt[0] = mfma32(w1, h3[0][1], mfma32(w0, h3[0][0], bias));
			nxt[1] = mfma32(w1, h3[1][1], mfma32(w0, h3[1][0], bias));
			const int u = t + 2 < kOutP / 16 ? t + 2 : kOutP / 16 - 1;
			w0 = w4s[(2 * u) * 64 + lane];
			w1 = w4s[(2 * u + 1) * 64 + lane];
			bias = b4s[4 * u + g];
			#pragma unroll
			for (int r = 0; r < 2; ++r)
				#pragma unroll
				for (int i = 0; i < 4; ++i) {
					const unsigned low = low0 - 16u * (unsigned)t - (unsigned)i;
					const float key = __uint_as_float((__float_as_uint(acc[r][i]) & 0xfffffc00u) | low);
					if (K == 1) tk[r][0] = fmaxf(tk[r][0], key);
					else insert4(tk[r], key);
				}
			if (LOGITS)
				#pragma unroll
				for (int r = 0; r < 2; ++r) {
					const int ray = tile * 32 + r * 16 + col;
					if (ray < n)
						#pragma unroll
						for (int i = 0; i < 4; ++i)
							logits[(size_t)ray * kOutP + 16 * t + 4 * g + i] = acc[r][i];
				}
			acc[0] = nxt[0];
			acc[1] = nxt[1];
		}
		if (logdist || LOGITS) {
			const int t = kClusters / 16;
			#pragma unroll
			for (int r = 0; r < 2; ++r) {
				const int ray = tile * 32 + r * 16 + col;
				if (ray >= n) continue;
				if (logdist && g == 0) logdist[ray] = acc[r][0];
				if (LOGITS)
					#pragma unroll
					for (int i = 0; i < 4; ++i)
						logits[(size_t)ray * kOutP + 16 * t + 4 * g + i] = acc[r][i];
			}
		}
		// merge the four lanes of each ray (groups g = 0..3)
		#pragma unroll
		for (int r = 0; r < 2; ++r) {
			#pragma unroll
			for (int off = 16; off <= 32; off <<= 1) {
				if (K == 1) {
					tk[r][0] = fmaxf(tk[r][0], __shfl_xor(tk[r][0], off));
					continue;
				}
				float o[4];
				#pragma unroll
				for (int i = 0; i < 4; ++i) o[i] = __shfl_xor(tk[r][i], off);
				#pragma unroll
				for (int i = 0; i < 4; ++i) insert4(tk[r], o[i]);
			}
			const int ray = tile * 32 + r * 16 + col;
			if (top && ray < n) {
				int c = -1;
				if (K == 1) c = g == 0 ? key_cluster(tk[r][0]) : -1;
				else c = key_cluster(g == 0 ? tk[r][0] : g == 1 ? tk[r][1] : g == 2 ? tk[r][2] : tk[r][3]);
				top[(size_t)ray * kTop + g] = c;
			}
		}
	}
}

// counter-based uniform in [-1, 1)
__host__ inline float init_uniform(uint64_t seed, uint64_t i) {
	uint64_t z = seed * 0x9e3779b97f4a7c15ull + i + 1;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	z ^= z >> 31;
	return (float)(z >> 40) * (2.0f / 16777216.0f) - 1.0f;
}

struct State {
	int max_batch = 0;
	float *params = nullptr, *mom = nullptr, *vel = nullptr, *grads = nullptr;
	uint16_t *wb = nullptr;      // bf16 copy of params (weights used from it)
	uint16_t *x = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr;
	float *out = nullptr;
	uint16_t *dout = nullptr, *dh3 = nullptr, *dh2 = nullptr, *dh1 = nullptr;
	uint16_t *wpack = nullptr;   // permuted bf16 weights of the fused inference kernel
	int cus = 0;                 // compute units of the device the state lives on
	int n_fwd = 0;               // rows of the last forward pass
	std::vector<void *> allocs;
	template <typename T> int alloc(T *&p, size_t count) {
		void *q = nullptr;
		if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return RTA_E_INVALID;
		allocs.push_back(q);
		p = reinterpret_cast<T *>(q);
		return RTA_OK;
	}
	~State() {
		for (void *q : allocs) (void)hipFree(q);
	}
};

inline unsigned blocks_for(long long items, int per_block) {
	return (unsigned)((items + per_block - 1) / per_block);
}

// He-uniform weights on the used entries, zero biases and padding, zero
// Adam moments and gradients, the bf16 copy, and the device's CU count
int init_state(State *p, uint64_t seed) {
	std::vector<float> h(kParams, 0.0f);
	auto fill = [&](int off, int rows, int cols_p, int cols, int fan_in) {
		const float a = sqrtf(6.0f / (float)fan_in);
		for (int r = 0; r < rows; ++r)
			for (int c = 0; c < cols; ++c)
				h[off + r * cols_p + c] = a * init_uniform(seed, (uint64_t)(off + r * cols_p + c));
	};
	fill(oW1, kHid, kInP, kIn, kIn);
	fill(oW2, kHid, kHid, kHid, kHid);
	fill(oW3, kHid, kHid, kHid, kHid);
	fill(oW4, kOut, kHid, kHid, kHid);
	RTA_HIP_TRY(hipMemcpy(p->params, h.data(), kParams * 4, hipMemcpyHostToDevice));
	RTA_HIP_TRY(hipMemset(p->mom, 0, kParams * 4));
	RTA_HIP_TRY(hipMemset(p->vel, 0, kParams * 4));
	RTA_HIP_TRY(hipMemset(p->grads, 0, kParams * 4));
	hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, 0, p->wb,
	                   p->params);
	RTA_LAUNCH_CHECK("to_bf16_kernel");
	int dev = 0;
	RTA_HIP_TRY(hipGetDevice(&dev));
	RTA_HIP_TRY(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev));
	if (p->cus < 1) p->cus = 1;
	RTA_HIP_TRY(hipDeviceSynchronize());
	return RTA_OK;
}

}  // namespace pred
}  // namespace rta_hip

using namespace rta_hip;
using namespace rta_hip::pred;

extern "C" {

int rta_pred_param_count(int *count) {
	if (!count) return invalid("rta_pred_param_count: NULL pointer");
	*count = kParams;
	return RTA_OK;
}

int rta_pred_create(int max_batch, uint64_t seed, rta_pred *out) {
	if (!out || max_batch <= 0 || max_batch > (1 << 24))
		return invalid("rta_pred_create: max_batch in 1..2^24 and a non-NULL handle");
	State *p = new (std::nothrow) State;
	if (!p) return invalid("rta_pred_create: out of host memory");
	p->max_batch = max_batch;
	const size_t B = ((size_t)max_batch + 15) / 16 * 16;
	int rc = RTA_OK;
	rc |= p->alloc(p->params, kParams);
	rc |= p->alloc(p->mom, kParams);
	rc |= p->alloc(p->vel, kParams);
	rc |= p->alloc(p->grads, kParams);
	rc |= p->alloc(p->wb, kParams);
	rc |= p->alloc(p->x, B * kInP);
	rc |= p->alloc(p->h1, B * kHid);
	rc |= p->alloc(p->h2, B * kHid);
	rc |= p->alloc(p->h3, B * kHid);
	rc |= p->alloc(p->out, B * kOutP);
	rc |= p->alloc(p->dout, B * kOutP);
	rc |= p->alloc(p->dh3, B * kHid);
	rc |= p->alloc(p->dh2, B * kHid);
	rc |= p->alloc(p->dh1, B * kHid);
	rc |= p->alloc(p->wpack, (size_t)kPackRecords * 8);
	if (rc != RTA_OK) {
		delete p;
		set_error("rta_pred_create: device allocation failed");
		return RTA_E_INVALID;
	}
	const int rc_init = init_state(p, seed);
	if (rc_init != RTA_OK) {
		delete p;
		return rc_init;
	}
	*out = p;
	return RTA_OK;
}

int rta_pred_destroy(rta_pred h) {
	delete reinterpret_cast<State *>(h);
	return RTA_OK;
}

int rta_pred_get_params(rta_pred h, float *host) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !host) return invalid("rta_pred_get_params: NULL pointer");
	RTA_HIP_TRY(hipMemcpy(host, p->params, kParams * 4, hipMemcpyDeviceToHost));
	return RTA_OK;
}

int rta_pred_set_params(rta_pred h, const float *host) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !host) return invalid("rta_pred_set_params: NULL pointer");
	RTA_HIP_TRY(hipMemcpy(p->params, host, kParams * 4, hipMemcpyHostToDevice));
	hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, 0, p->wb,
	                   p->params);
	RTA_LAUNCH_CHECK("to_bf16_kernel");
	RTA_HIP_TRY(hipDeviceSynchronize());
	return RTA_OK;
}

int rta_pred_get_grads(rta_pred h, float *host) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !host) return invalid("rta_pred_get_grads: NULL pointer");
	RTA_HIP_TRY(hipMemcpy(host, p->grads, kParams * 4, hipMemcpyDeviceToHost));
	return RTA_OK;
}

int rta_pred_zero_grads(rta_pred h, rta_stream s) {
	State *p = reinterpret_cast<State *>(h);
	if (!p) return invalid("rta_pred_zero_grads: NULL handle");
	return rta_hip_fill_u32(reinterpret_cast<uint32_t *>(p->grads), 0u, kParams, s);
}

int rta_pred_forward(rta_pred h, const float *o, const float *d, int n, const float *scene_min,
                     const float *scene_max, float *out, int32_t *top, float *logdist,
                     rta_stream s) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !o || !d || !scene_min || !scene_max) return invalid("rta_pred_forward: NULL pointer");
	if (n <= 0 || n > p->max_batch) return invalid("rta_pred_forward: n must be in 1..max_batch");
	hipStream_t st = as_stream(s);
	float3 lo = make_float3(scene_min[0], scene_min[1], scene_min[2]);
	float3 sc = make_float3(2.0f / (scene_max[0] - scene_min[0]), 2.0f / (scene_max[1] - scene_min[1]),
	                        2.0f / (scene_max[2] - scene_min[2]));
	hipLaunchKernelGGL(encode_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, p->x, o, d, n, lo,
	                   sc);
	RTA_LAUNCH_CHECK("encode_kernel");
	const long long rt = (n + 15) / 16;
	hipLaunchKernelGGL(fwd_kernel<false>, dim3(blocks_for(rt * (kHid / 16), 4)), dim3(256), 0, st,
	                   (void *)p->h1, kHid, p->x, kInP, p->wb + oW1, p->params + oB1, kInP, kHid, n);
	hipLaunchKernelGGL(fwd_kernel<false>, dim3(blocks_for(rt * (kHid / 16), 4)), dim3(256), 0, st,
	                   (void *)p->h2, kHid, p->h1, kHid, p->wb + oW2, p->params + oB2, kHid, kHid, n);
	hipLaunchKernelGGL(fwd_kernel<false>, dim3(blocks_for(rt * (kHid / 16), 4)), dim3(256), 0, st,
	                   (void *)p->h3, kHid, p->h2, kHid, p->wb + oW3, p->params + oB3, kHid, kHid, n);
	hipLaunchKernelGGL(fwd_kernel<true>, dim3(blocks_for(rt * (kOutP / 16), 4)), dim3(256), 0, st,
	                   (void *)p->out, kOutP, p->h3, kHid, p->wb + oW4, p->params + oB4, kHid, kOutP,
	                   n);
	RTA_LAUNCH_CHECK("fwd_kernel");
	if (out)
		RTA_HIP_TRY(hipMemcpyAsync(out, p->out, (size_t)n * kOutP * 4, hipMemcpyDeviceToDevice, st));
	if (top || logdist) {
		hipLaunchKernelGGL(topk_kernel, dim3(blocks_for(n, 4)), dim3(256), 0, st, top, logdist,
		                   p->out, n);
		RTA_LAUNCH_CHECK("topk_kernel");
	}
	p->n_fwd = n;
	return RTA_OK;
}

int rta_pred_infer(rta_pred h, const float *o, const float *d, int n, const float *scene_min,
                   const float *scene_max, int32_t *top, int top_k, float *logdist,
                   float *logits, rta_stream s) {
	RTA_FAULT_POINT("rta_pred_infer");
	State *p = reinterpret_cast<State *>(h);
	if (!p || !o || !d || !scene_min || !scene_max) return invalid("rta_pred_infer: NULL pointer");
	if (!top && !logdist && !logits) return invalid("rta_pred_infer: no output requested");
	if (n < 0) return invalid("rta_pred_infer: n < 0");
	if (top_k != 1 && top_k != kTop) return invalid("rta_pred_infer: top_k must be 1 or 4");
	if (n == 0) return RTA_OK;
	hipStream_t st = as_stream(s);
	// the permuted weights follow the current parameters (one small launch)
	hipLaunchKernelGGL(pack_infer_kernel, dim3(blocks_for((long long)kPackRecords * 8, 256)),
	                   dim3(256), 0, st, p->wpack, p->wb);
	RTA_LAUNCH_CHECK("pack_infer_kernel");
	float3 lo = make_float3(scene_min[0], scene_min[1], scene_min[2]);
	float3 sc = make_float3(2.0f / (scene_max[0] - scene_min[0]), 2.0f / (scene_max[1] - scene_min[1]),
	                        2.0f / (scene_max[2] - scene_min[2]));
	const long long tiles = ((long long)n + 31) / 32;
	const long long need = (tiles + kInferWaves - 1) / kInferWaves;
	const unsigned blocks = (unsigned)(need < p->cus ? need : p->cus);
	const v4u *wp = reinterpret_cast<const v4u *>(p->wpack);
	const dim3 grid(blocks), block(kInferWaves * kWave);
	if (logits && top_k == 1)
		hipLaunchKernelGGL((infer_kernel<true, 1>), grid, block, 0, st, wp, p->params, o, d, n, lo,
		                   sc, top, logdist, logits);
	else if (logits)
		hipLaunchKernelGGL((infer_kernel<true, kTop>), grid, block, 0, st, wp, p->params, o, d, n,
		                   lo, sc, top, logdist, logits);
	else if (top_k == 1)
		hipLaunchKernelGGL((infer_kernel<false, 1>), grid, block, 0, st, wp, p->params, o, d, n, lo,
		                   sc, top, logdist, nullptr);
	else
		hipLaunchKernelGGL((infer_kernel<false, kTop>), grid, block, 0, st, wp, p->params, o, d, n,
		                   lo, sc, top, logdist, nullptr);
	RTA_LAUNCH_CHECK("infer_kernel");
	return RTA_OK;
}

int rta_pred_backward(rta_pred h, const int32_t *label, const float *target, int n, float *loss,
                      rta_stream s) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !label || !target || !loss) return invalid("rta_pred_backward: NULL pointer");
	if (n != p->n_fwd) return invalid("rta_pred_backward: n differs from the last forward pass");
	hipStream_t st = as_stream(s);
	hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(n, 4)), dim3(256), 0, st, p->dout, p->out,
	                   label, target, n, loss);
	RTA_LAUNCH_CHECK("loss_kernel");
	const long long rt = (n + 15) / 16;
	const int chunks = (n + kRowChunk - 1) / kRowChunk;
	float *g = p->grads;
	// layer 4: dW4, db4, dh3
	hipLaunchKernelGGL(bwd_weight_kernel, dim3(blocks_for((long long)(kOutP / 16) * (kHid / 16) * chunks, 4)),
	                   dim3(256), 0, st, g + oW4, p->dout, kOutP, p->h3, kHid, kHid, kOutP, n);
	hipLaunchKernelGGL(bwd_bias_kernel, dim3(blocks_for(kOutP, kWave), chunks), dim3(256), 0, st,
	                   g + oB4, p->dout, kOutP, kOutP, n);
	hipLaunchKernelGGL(bwd_data_kernel, dim3(blocks_for(rt * (kHid / 16), 4)), dim3(256), 0, st,
	                   p->dh3, p->dout, kOutP, p->wb + oW4, kHid, kOutP, p->h3, n);
	// layer 3
	hipLaunchKernelGGL(bwd_weight_kernel, dim3(blocks_for((long long)(kHid / 16) * (kHid / 16) * chunks, 4)),
	                   dim3(256), 0, st, g + oW3, p->dh3, kHid, p->h2, kHid, kHid, kHid, n);
	hipLaunchKernelGGL(bwd_bias_kernel, dim3(blocks_for(kHid, kWave), chunks), dim3(256), 0, st,
	                   g + oB3, p->dh3, kHid, kHid, n);
	hipLaunchKernelGGL(bwd_data_kernel, dim3(blocks_for(rt * (kHid / 16), 4)), dim3(256), 0, st,
	                   p->dh2, p->dh3, kHid, p->wb + oW3, kHid, kHid, p->h2, n);
	// layer 2
	hipLaunchKernelGGL(bwd_weight_kernel, dim3(blocks_for((long long)(kHid / 16) * (kHid / 16) * chunks, 4)),
	                   dim3(256), 0, st, g + oW2, p->dh2, kHid, p->h1, kHid, kHid, kHid, n);
	hipLaunchKernelGGL(bwd_bias_kernel, dim3(blocks_for(kHid, kWave), chunks), dim3(256), 0, st,
	                   g + oB2, p->dh2, kHid, kHid, n);
	hipLaunchKernelGGL(bwd_data_kernel, dim3(blocks_for(rt * (kHid / 16), 4)), dim3(256), 0, st,
	                   p->dh1, p->dh2, kHid, p->wb + oW2, kHid, kHid, p->h1, n);
	// layer 1 (no input gradient)
	hipLaunchKernelGGL(bwd_weight_kernel, dim3(blocks_for((long long)(kHid / 16) * (kInP / 16) * chunks, 4)),
	                   dim3(256), 0, st, g + oW1, p->dh1, kHid, p->x, kInP, kInP, kHid, n);
	hipLaunchKernelGGL(bwd_bias_kernel, dim3(blocks_for(kHid, kWave), chunks), dim3(256), 0, st,
	                   g + oB1, p->dh1, kHid, kHid, n);
	RTA_LAUNCH_CHECK("backward kernels");
	return RTA_OK;
}

int rta_pred_adam(rta_pred h, float lr, int step, rta_stream s) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || step < 1) return invalid("rta_pred_adam: NULL handle or step < 1");
	const float c1 = 1.0f / (1.0f - powf(0.9f, (float)step));
	const float c2 = 1.0f / (1.0f - powf(0.999f, (float)step));
	hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, as_stream(s),
	                   p->params, p->mom, p->vel, p->grads, p->wb, lr, c1, c2, nullptr);
	RTA_LAUNCH_CHECK("adam_kernel");
	return RTA_OK;
}

int rta_pred_adam_device(rta_pred h, float lr, uint32_t *step, rta_stream s) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !step) return invalid("rta_pred_adam_device: NULL argument");
	const int rc = rta_hip_add_u32(step, 1u, s);
	if (rc != RTA_OK) return rc;
	hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, as_stream(s),
	                   p->params, p->mom, p->vel, p->grads, p->wb, lr, 0.0f, 0.0f, step);
	RTA_LAUNCH_CHECK("adam_kernel");
	return RTA_OK;
}

int rta_pred_sample(const float *o, const float *d, const rta_hit *hits, const uint32_t *count,
                    int n_out, const uint32_t *frame_word, uint64_t seed, float *o_out,
                    float *d_out, rta_hit *hits_out, rta_stream s) {
	if (!o || !d || !hits || !count || !o_out || !d_out || !hits_out || n_out <= 0)
		return invalid("rta_pred_sample: bad arguments");
	hipLaunchKernelGGL(sample_kernel, dim3(blocks_for(n_out, 256)), dim3(256), 0, as_stream(s), o,
	                   d, hits, count, n_out, frame_word, seed, o_out, d_out, hits_out);
	RTA_LAUNCH_CHECK("sample_kernel");
	return RTA_OK;
}

}  // extern "C"
