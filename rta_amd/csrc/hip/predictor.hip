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
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2b __attribute__((ext_vector_type(2)));

// ------------------------------------------------- MFMA weight packs
//
// The fused kernels compute every product transposed, D = W H (rows =
// output features, columns = rays), with the 16x16x32 bf16 MFMA: lane l
// supplies A row l%16 and B column (ray) l%16, and the 8 k-slots of lane
// group g = l/16.  Lane l then holds D[4g+i][l%16], which is already the
// next product's B operand once the k-slots of a 32-wide k-step s are
//   slot j<4  -> feature 32s + 4g + j        (tile 2s)
//   slot j>=4 -> feature 32s + 16 + 4g + j-4 (tile 2s+1)
// The first layer's slots are the encoded inputs rotated by 6 (k' =
// feature - 6 for the 96 sin/cos features, then the 6 raw inputs at k' =
// 96..101, zeros from 102) and assigned k' = 32s + 8g + j: slot (s, g, j) is
// octave block 4s+g, which a lane computes by itself.
//
// Forward pack: W as A operands, one 16-B record per (16-row tile, k-step,
// lane).  Backward pack: W^T as A operands of dH = W^T dZ (rows = input
// features k, k-steps over output features n, same slot rule), for layers
// 4, 3, 2.  Every writer of the parameters (initialisation, set_params,
// Adam) rewrites the bf16 copy and both packs, so they always match.
constexpr int kRec1 = 4 * 4 * 64;                     // layer 1: 4 tiles x 4 k-steps
constexpr int kRecH = 4 * 2 * 64;                     // layers 2, 3
constexpr int kRec4 = (kOutP / 16) * 2 * 64;          // layer 4: 65 tiles x 2 k-steps
constexpr int pW1 = 0, pW2 = pW1 + kRec1, pW3 = pW2 + kRecH, pW4 = pW3 + kRecH;
constexpr int kPackRecords = pW4 + kRec4;             // 16-B records (165,888 B)
constexpr int kStepsT4 = (kOutP + 31) / 32;           // 33 k-steps over W4's rows
constexpr int qW4 = 0, qW3 = qW4 + 4 * kStepsT4 * 64, qW2 = qW3 + kRecH;
constexpr int kPackTRecords = qW2 + kRecH;            // 9,472 records (151,552 B)

// group and slot of feature f within its 32-wide k-step (hidden layers)
__device__ __host__ inline void slot_of(int f, int &g, int &j) {
	const int r = f & 15;
	g = r >> 2;
	j = 4 * ((f >> 4) & 1) + (r & 3);
}

// element (bf16 index) of the forward pack holding W_L[n][k]; -1 for W1's
// padding columns, which are not packed
__device__ __host__ inline int fwd_elem(int L, int n, int k) {
	if (L == 1) {
		if (k >= kIn) return -1;
		const int kp = k >= 6 ? k - 6 : k + 96;
		return 8 * (pW1 + ((n >> 4) * 4 + (kp >> 5)) * 64 + 16 * ((kp & 31) >> 3) + (n & 15)) +
		       (kp & 7);
	}
	const int base = L == 2 ? pW2 : (L == 3 ? pW3 : pW4);
	int g, j;
	slot_of(k, g, j);
	return 8 * (base + ((n >> 4) * 2 + (k >> 5)) * 64 + 16 * g + (n & 15)) + j;
}

// element of the backward pack holding W_L[n][k], L = 2..4
__device__ __host__ inline int bwd_elem(int L, int n, int k) {
	const int base = L == 4 ? qW4 : (L == 3 ? qW3 : qW2);
	const int steps = L == 4 ? kStepsT4 : 2;
	int g, j;
	slot_of(n, g, j);
	return 8 * (base + ((k >> 4) * steps + (n >> 5)) * 64 + 16 * g + (k & 15)) + j;
}

// layer, row and column of parameter i; false for a bias
__device__ __host__ inline bool weight_at(int i, int &L, int &n, int &k) {
	int off, ld;
	if (i < oB1) { L = 1; off = oW1; ld = kInP; }
	else if (i >= oW2 && i < oB2) { L = 2; off = oW2; ld = kHid; }
	else if (i >= oW3 && i < oB3) { L = 3; off = oW3; ld = kHid; }
	else if (i >= oW4 && i < oB4) { L = 4; off = oW4; ld = kHid; }
	else return false;
	n = (i - off) / ld;
	k = (i - off) % ld;
	return true;
}

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

// The bf16 copy and both MFMA packs of parameter i (biases: the copy only).
__device__ inline void store_param(int i, float v, uint16_t *__restrict__ wb,
                                   uint16_t *__restrict__ wp, uint16_t *__restrict__ wpt) {
	const uint16_t b = to_bf16(v);
	wb[i] = b;
	int L, n, k;
	if (!weight_at(i, L, n, k)) return;
	const int e = fwd_elem(L, n, k);
	if (e >= 0) wp[e] = b;
	if (L >= 2) wpt[bwd_elem(L, n, k)] = b;
}

// One Adam step of parameter i with gradient gi (beta 0.9 / 0.999, eps
// 1e-8; c1, c2: the bias corrections); returns the new parameter.
__device__ inline float adam_step(float *__restrict__ p, float *__restrict__ mo,
                                  float *__restrict__ ve, int i, float gi, float lr, float c1,
                                  float c2) {
	const float m1 = 0.9f * mo[i] + 0.1f * gi;
	const float v1 = 0.999f * ve[i] + 0.001f * gi * gi;
	mo[i] = m1;
	ve[i] = v1;
	const float pi = p[i] - lr * (m1 * c1) / (sqrtf(v1 * c2) + 1e-8f);
	p[i] = pi;
	return pi;
}

__device__ inline void bias_corrections(const uint32_t *step_word, float &c1, float &c2) {
	const float st = (float)*step_word;
	c1 = 1.0f / (1.0f - powf(0.9f, st));
	c2 = 1.0f / (1.0f - powf(0.999f, st));
}

// Adam on the packed parameters with the accumulated gradients g.
// step_word (optional): take the step from device memory (bias corrections
// computed here) instead of c1, c2, so a captured launch advances with it.
__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, float *__restrict__ mo,
                                                   float *__restrict__ ve,
                                                   const float *__restrict__ g,
                                                   uint16_t *__restrict__ wb,
                                                   uint16_t *__restrict__ wp,
                                                   uint16_t *__restrict__ wpt, float lr, float c1,
                                                   float c2, const uint32_t *__restrict__ step_word) {
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kParams) return;
	if (step_word) bias_corrections(step_word, c1, c2);
	store_param(i, adam_step(p, mo, ve, i, g[i], lr, c1, c2), wb, wp, wpt);
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

// the bf16 copy and packs of every parameter (after initialisation or
// set_params)
__global__ __launch_bounds__(256) void store_params_kernel(uint16_t *__restrict__ wb,
                                                           uint16_t *__restrict__ wp,
                                                           uint16_t *__restrict__ wpt,
                                                           const float *__restrict__ p) {
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i < kParams) store_param(i, p[i], wb, wp, wpt);
}

// ------------------------------------------------- fused inference
//
// One kernel from rays to top-4 treelets: activations stay in registers and
// the 1040 logits of a ray never reach memory (operand layout and weight
// packs: see the MFMA weight packs above).
//
// Top-4 per ray: each logit becomes a key whose low 10 mantissa bits are
// replaced by 1023 - treelet (distinct keys; the order of the keys is the
// order of the logits truncated to 13 mantissa bits, ties to the lower
// treelet for positive logits), kept in a sorted register quadruple by one
// max and three med3 per key, and the four lanes of a ray merged by
// shuffles at the end.

constexpr int kInferWaves = 16;                       // 1024 threads per block

__device__ inline unsigned bf16x2(float a, float b) {
	return __builtin_bit_cast(unsigned, __builtin_convertvector((v2f){a, b}, v2b));
}

__device__ inline v4f mfma32(v4u a, v4u b, v4f c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8b, a),
	                                               __builtin_bit_cast(v8b, b), c, 0, 0, 0);
}

// one 64-wide ReLU layer for NS sets of 16 rays; in: KS k-steps of B
// operands per set, out: the next layer's 2 k-steps per set
template <int KS, int NS>
__device__ inline void hidden_layer(v4u out[NS][2], const v4u in[NS][KS],
                                    const v4u *__restrict__ wp, const float *__restrict__ bias,
                                    int lane) {
	unsigned d[NS][4][2];
	#pragma unroll
	for (int t = 0; t < 4; ++t) {
		const float4 b = *reinterpret_cast<const float4 *>(bias + 16 * t + 4 * (lane >> 4));
		v4f a[NS];
		#pragma unroll
		for (int r = 0; r < NS; ++r) a[r] = (v4f){b.x, b.y, b.z, b.w};
		#pragma unroll
		for (int s = 0; s < KS; ++s) {
			const v4u w = wp[(t * KS + s) * 64 + lane];
			#pragma unroll
			for (int r = 0; r < NS; ++r) a[r] = mfma32(w, in[r][s], a[r]);
		}
		#pragma unroll
		for (int r = 0; r < NS; ++r) {
			d[r][t][0] = bf16x2(fmaxf(a[r][0], 0.f), fmaxf(a[r][1], 0.f));
			d[r][t][1] = bf16x2(fmaxf(a[r][2], 0.f), fmaxf(a[r][3], 0.f));
		}
	}
	#pragma unroll
	for (int r = 0; r < NS; ++r)
		#pragma unroll
		for (int s = 0; s < 2; ++s)
			out[r][s] = (v4u){d[r][2 * s][0], d[r][2 * s][1], d[r][2 * s + 1][0], d[r][2 * s + 1][1]};
}

// layer-1 B operands of the ray in lane column col (zero beyond n): octave
// blocks 4s+g for s = 0..2 and the raw inputs (group 0) at s = 3
__device__ inline void encode_operands(v4u x[4], const float *__restrict__ ro,
                                       const float *__restrict__ rd, int ray, int n, float3 lo,
                                       float3 scale, int g) {
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
		x[s] = (v4u){bf16x2(e[0], e[1]), bf16x2(e[2], e[3]), bf16x2(e[4], e[5]),
		             bf16x2(e[6], e[7])};
	}
	x[3] = g == 0 ? (v4u){bf16x2(v0, v1), bf16x2(v2, v3), bf16x2(v4, v5), 0u} : (v4u){0u, 0u, 0u, 0u};
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
		v4u x[2][4];
		#pragma unroll
		for (int r = 0; r < 2; ++r) encode_operands(x[r], ro, rd, tile * 32 + r * 16 + col, n, lo, scale, g);
		// the hidden layers' weights and biases are loaded per tile (from the
		// caches): the empty asm hides their loop invariance, which would
		// otherwise hoist all of them into registers and spill (an opaque
		// offset, so that the loads stay global rather than flat)
		int z = 0;
		asm volatile("" : "+s"(z));
		const v4u *wq = wp + z;
		const float *pq = params + z;
		v4u h1[2][2], h2[2][2], h3[2][2];
		hidden_layer<4, 2>(h1, x, wq + pW1, pq + oB1, lane);
		hidden_layer<2, 2>(h2, h1, wq + pW2, pq + oB2, lane);
		hidden_layer<2, 2>(h3, h2, wq + pW3, pq + oB3, lane);
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

// --------------------------------------------- fused training step
//
// Three launches per batch of m rays, without atomics (the results do not
// depend on scheduling):
//  1. train_rows_kernel, per set of 16 rays and wave: the forward pass as in
//     the fused inference; the 1040 outputs from the LDS-resident W4 pack
//     twice, never stored -- first for the running max and sum of exp of
//     the 1024 treelet logits, then tile by tile for dOut = (softmax -
//     onehot(label)) / m and 2 (logdist - target) / m, which is at once the
//     B operand of dH3 = W4^T dOut; then dZ3 = dH3 [h3 > 0], dZ2, dZ1
//     through the backward pack.  It stores what the weight gradients need,
//     feature-major bf16 rows ([feature][ray], ld rays per row): the layer
//     inputs x (rows in operand order k'), h1, h2, h3 and the output
//     gradients dOut, dZ3, dZ2, dZ1; and one partial loss per set.
//  2. grad_partials_kernel: dW = dZ X^T and db = dZ 1 of every layer, MFMA
//     products over the rays in splits of 256 (rows padded to whole splits
//     by rays without a hit); the kGradWaves splits of a block are added in
//     LDS (in wave order) and stored as one partial.
//  3. train_adam_kernel: per parameter the sum of its partials (in order),
//     Adam, the bf16 copy and packs; block 0 adds the sets' partial losses
//     (in a fixed order) to the loss word.

constexpr int kTrainSets = 2;              // 16-ray sets per block of train_rows_kernel (2 waves each)

// The rows the weight gradients read are stored in MFMA operand order: for
// each tile of 16 features and each k-step of 32 rays, the 16 x 32 bf16
// block as one 1-KiB record, lane l's 8 values (feature l%16, rays 8(l/16)
// ..+7 of the k-step) at 16 l.  Element (f, r) of an array with ld rays per
// row (ld a multiple of 32) is at rec_offset(f, r, ld).
struct TrainRows {
	uint16_t *x, *h1, *h2, *h3;            // features 112 (operand order k'), 64, 64, 64
	uint16_t *dout, *dz3, *dz2, *dz1;      // features 1040, 64, 64, 64
	float *loss_part;                      // two per 16-ray set (one per half)
	int ld;                                // rays per feature (a multiple of 256)
};

__device__ __host__ inline size_t rec_offset(int f, int r, int ld) {
	return ((size_t)(f >> 4) * (ld >> 5) + (r >> 5)) * 512 +
	       (size_t)(16 * ((r & 31) >> 3) + (f & 15)) * 8 + (r & 7);
}

__device__ inline uint16_t slot_bits(v4u v, int j) {
	const unsigned w = v[j >> 1];
	return (uint16_t)((j & 1) ? (w >> 16) : (w & 0xffffu));
}

// Stores go through a per-wave LDS tile of 32 features x 16 rays (bf16,
// row stride kTRow): each lane writes its 8 slots, then reads 8 consecutive
// rays of one feature, which are 16 contiguous bytes of a record; the 64
// lanes write two 512-B runs.  LDS rows are 48 B: 16-B aligned, and the four
// lane groups of a hidden-layer write (rows 4g + j) fall in distinct banks.
// A wave's LDS accesses execute in order, so no barrier is needed.
constexpr int kTRow = 24;
constexpr int kTTile = 32 * kTRow;          // elements per wave

// one 32-feature k-step operand v of the set's rays (slots: layer1 ? k' =
// 32s + 8g + j : the hidden-layer rule) as features f0..f0+31 of dst
// (features >= nrows skipped), rays ray0..ray0+15 (ray0 a multiple of 16)
template <bool STREAM = false>
__device__ inline void store_kstep(uint16_t *__restrict__ dst, int ld, int ray0, int f0,
                                   int nrows, v4u v, bool layer1, uint16_t *t, int lane) {
	const int g = lane >> 4, c = lane & 15;
	#pragma unroll
	for (int j = 0; j < 8; ++j) {
		const int fl = layer1 ? 8 * g + j : (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
		t[fl * kTRow + c] = slot_bits(v, j);
	}
	const int fr = lane >> 1, hh = lane & 1;
	const v4u row = *reinterpret_cast<const v4u *>(t + fr * kTRow + 8 * hh);
	if (f0 + fr < nrows) {
		v4u *q = reinterpret_cast<v4u *>(dst + rec_offset(f0 + fr, ray0 + 8 * hh, ld));
		if (STREAM) __builtin_nontemporal_store(row, q);   // read once, by the next kernel
		else *q = row;
	}
}

// the 2 hidden-layer k-steps as features 0..63
__device__ inline void store_hidden(uint16_t *__restrict__ dst, int ld, int ray0, const v4u h[2],
                                    uint16_t *t, int lane) {
	store_kstep(dst, ld, ray0, 0, kHid, h[0], false, t, lane);
	store_kstep(dst, ld, ray0, 32, kHid, h[1], false, t, lane);
}

// dZ = dH [h > 0] as B operands: dh[kt] holds features 16kt + 4g + i, h is
// the layer's output in operand layout (a bf16 is > 0 iff, read as a signed
// 16-bit integer, it is > 0)
__device__ inline void relu_back(v4u dz[2], const v4f dh[4], const v4u h[2]) {
	#pragma unroll
	for (int s = 0; s < 2; ++s) {
		unsigned w[4];
		#pragma unroll
		for (int half = 0; half < 2; ++half) {
			const v4f a = dh[2 * s + half];
			#pragma unroll
			for (int q = 0; q < 2; ++q) {
				const unsigned hw = h[s][2 * half + q];
				const float d0 = (short)(hw & 0xffffu) > 0 ? a[2 * q] : 0.f;
				const float d1 = (short)(hw >> 16) > 0 ? a[2 * q + 1] : 0.f;
				w[2 * half + q] = bf16x2(d0, d1);
			}
		}
		dz[s] = (v4u){w[0], w[1], w[2], w[3]};
	}
}

// dH = W^T dZ of a 64-wide layer from the lane's 8 backward-pack records
__device__ inline void data_back(v4f dh[4], const v4u dz[2], const v4u w[8]) {
	#pragma unroll
	for (int kt = 0; kt < 4; ++kt) {
		v4f a = {0.f, 0.f, 0.f, 0.f};
		#pragma unroll
		for (int s = 0; s < 2; ++s) a = mfma32(w[kt * 2 + s], dz[s], a);
		dh[kt] = a;
	}
}

// log2(e): exp(a - m) = exp2(fma(a, kLog2e, -m kLog2e)) on the hardware exp2
constexpr float kLog2e = 1.44269504088896341f;

__device__ inline float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }

// Two waves per set of 16 rays: both run the forward pass; wave half h
// takes logit tiles 32h..32h+31 (k-steps 16h..16h+15) of the softmax passes;
// the halves' (max, sum) per ray and the second half's dH3 are exchanged in
// LDS.  Half 0 then runs the hidden layers' data gradients, half 1 stores
// the layer inputs and the log distance's k-step.  SETS sets per block.
template <int SETS>
__global__ __launch_bounds__(2 * SETS * kWave) void train_rows_kernel(
    const v4u *__restrict__ wp, const v4u *__restrict__ wpt, const float *__restrict__ params,
    const float *__restrict__ ro, const float *__restrict__ rd, const int32_t *__restrict__ label,
    const float *__restrict__ target, int n, int n_sets, float3 lo, float3 scale, float inv_m,
    TrainRows out, uint32_t *__restrict__ step) {
	// LDS: the W4 pack and bias, each wave's transpose tile and the halves'
	// exchange (152,128 B with SETS = 2)
	__shared__ v4u w4s[kRec4];
	__shared__ v4f b4s[kOutP / 4];
	__shared__ __attribute__((aligned(16))) uint16_t tr[2 * SETS][kTTile];
	__shared__ v4f xch[SETS][4][kWave];
	__shared__ float2 stat[SETS][2][16];
	const int lane = __lane_id(), g = lane >> 4, col = lane & 15;
	const int wave = (int)(threadIdx.x / kWave), sl = wave >> 1, half = wave & 1;
	// W4's pack goes to LDS by direct global->LDS loads (64 records of 16 B
	// per instruction, lane l's record at base + 16 l), all in flight while
	// the first set's forward pass runs; waited for before its pass 1
	for (int i = wave; i < kRec4 / 64; i += 2 * SETS)
		__builtin_amdgcn_global_load_lds(
		    (const __attribute__((address_space(1))) void *)(wp + pW4 + 64 * i + lane),
		    (__attribute__((address_space(3))) void *)(w4s + 64 * i), 16, 0, 0);
	for (int i = threadIdx.x; i < kOutP / 4; i += 2 * SETS * kWave)
		b4s[i] = reinterpret_cast<const v4f *>(params + oB4)[i];
	if (blockIdx.x == 0 && threadIdx.x == 0) *step += 1u;
	uint16_t *t = tr[wave];
	const int ld = out.ld;
	// the loop runs the same number of times in every wave of the block, at
	// least once (its barriers are reached by all of them)
	for (int base = blockIdx.x * SETS; base < n_sets; base += gridDim.x * SETS) {
		const int set = base + sl;
		const bool act = set < n_sets;
		const int ray0 = set * 16, ray = ray0 + col;
		const int lab = act && ray < n ? label[ray] : -1;      // >= 0: a hit, in the loss
		const float tgt = act && ray < n ? target[ray] : 0.f;
		// per-set reads of the hidden layers' records (an opaque zero offset
		// keeps them from being hoisted out of the set loop and held in
		// registers throughout)
		int z = 0;
		asm volatile("" : "+s"(z));
		const v4u *wq = wp + z;
		const v4u *wqt = wpt + z;
		v4u xs[1][4];
		encode_operands(xs[0], ro, rd, act ? ray : n, n, lo, scale, g);
		const float *pq = params + z;
		v4u h1[1][2], h2[1][2], h3[1][2];
		hidden_layer<4, 1>(h1, xs, wq + pW1, pq + oB1, lane);
		hidden_layer<2, 1>(h2, h1, wq + pW2, pq + oB2, lane);
		hidden_layer<2, 1>(h3, h2, wq + pW3, pq + oB3, lane);
		if (base == blockIdx.x * SETS) {                  // W4 and its bias in LDS
			__builtin_amdgcn_s_waitcnt(0);
			__syncthreads();
		}

		// pass 1 over the half's tiles u0..u0+31: running max and sum of
		// exp.  Software-pipelined: the products of tile u+1 are in flight
		// while tile u's are reduced, and the records of tile u+2 are read
		// meanwhile; two steps per iteration with the roles of the register
		// sets swapped, so no value is copied across iterations (a copy
		// would wait for its load or product).
		const int u0 = 32 * half;
		float mrun = -__builtin_inff(), srun = 0.f;
		auto p1_step = [&](int u, const v4f &cur, v4f &nxt, const v4u &ua, const v4u &ub,
		                   const v4f &ubias, v4u &la, v4u &lb, v4f &lbias) {
			nxt = mfma32(ub, h3[0][1], mfma32(ua, h3[0][0], ubias));
			const int q = u + 2 <= kClusters / 16 ? u + 2 : kClusters / 16;
			la = w4s[(2 * q) * 64 + lane];
			lb = w4s[(2 * q + 1) * 64 + lane];
			lbias = b4s[4 * q + g];
			const float nm = fmaxf(mrun, fmaxf(fmaxf(cur[0], cur[1]), fmaxf(cur[2], cur[3])));
			const float nml = nm * kLog2e;
			const float e = ((exp2_hw(__builtin_fmaf(cur[0], kLog2e, -nml)) +
			                  exp2_hw(__builtin_fmaf(cur[1], kLog2e, -nml))) +
			                 (exp2_hw(__builtin_fmaf(cur[2], kLog2e, -nml)) +
			                  exp2_hw(__builtin_fmaf(cur[3], kLog2e, -nml))));
			srun = __builtin_fmaf(srun, exp2_hw(__builtin_fmaf(mrun, kLog2e, -nml)), e);
			mrun = nm;
		};
		v4f k0 = mfma32(w4s[(2 * u0 + 1) * 64 + lane], h3[0][1],
		                mfma32(w4s[(2 * u0) * 64 + lane], h3[0][0], b4s[4 * u0 + g])), k1;
		v4u a0 = w4s[(2 * u0 + 2) * 64 + lane], b0 = w4s[(2 * u0 + 3) * 64 + lane], a1, b1;
		v4f i0 = b4s[4 * u0 + 4 + g], i1;
		for (int u = u0; u < u0 + kClusters / 32; u += 2) {
			p1_step(u, k0, k1, a0, b0, i0, a1, b1, i1);
			p1_step(u + 1, k1, k0, a1, b1, i1, a0, b0, i0);
		}
		const float dist = k0[0];           // half 1: tile 64, feature 1024 in group 0
		float mh = fmaxf(mrun, __shfl_xor(mrun, 16));
		mh = fmaxf(mh, __shfl_xor(mh, 32));
		float sh = srun * exp2_hw(__builtin_fmaf(mrun, kLog2e, -mh * kLog2e));
		sh += __shfl_xor(sh, 16);
		sh += __shfl_xor(sh, 32);
		if (g == 0) stat[sl][half][col] = make_float2(mh, sh);
		__syncthreads();
		const float2 st0 = stat[sl][0][col], st1 = stat[sl][1][col];
		const float mx = fmaxf(st0.x, st1.x);
		const float mxl = mx * kLog2e;
		const float se = st0.y * exp2_hw(__builtin_fmaf(st0.x, kLog2e, -mxl)) +
		                 st1.y * exp2_hw(__builtin_fmaf(st1.x, kLog2e, -mxl));
		const float log_z = __logf(se) + mx;
		// dOut = e * scale - [f == label] / m (zero for rays without a hit)
		const float scale_g = lab >= 0 ? inv_m / se : 0.f;

		// pass 2 over the half's k-steps s0..s0+15: dOut per pair of tiles,
		// stored and multiplied into (the half's part of) dH3 = W4^T dOut;
		// products one k-step ahead, LDS records two, backward records one;
		// pipelined as pass 1
		const int s0 = 16 * half;
		v4f dh[4];
		#pragma unroll
		for (int kt = 0; kt < 4; ++kt) dh[kt] = (v4f){0.f, 0.f, 0.f, 0.f};
		float vlab = 0.f;
		struct Rec { v4u r0, r1, r2, r3; v4f c0, c1; };
		auto p2_step = [&](int s, const v4f &pa, const v4f &pb, v4f &na, v4f &nb, const Rec &ru,
		                   Rec &rl, const v4u bu[4], v4u bl[4]) {
			na = mfma32(ru.r1, h3[0][1], mfma32(ru.r0, h3[0][0], ru.c0));
			nb = mfma32(ru.r3, h3[0][1], mfma32(ru.r2, h3[0][0], ru.c1));
			const int q = s + 2 < kClusters / 32 ? s + 2 : kClusters / 32 - 1;
			rl.r0 = w4s[(4 * q) * 64 + lane];
			rl.r1 = w4s[(4 * q + 1) * 64 + lane];
			rl.r2 = w4s[(4 * q + 2) * 64 + lane];
			rl.r3 = w4s[(4 * q + 3) * 64 + lane];
			rl.c0 = b4s[8 * q + g];
			rl.c1 = b4s[8 * q + 4 + g];
			#pragma unroll
			for (int kt = 0; kt < 4; ++kt) bl[kt] = wqt[qW4 + (kt * kStepsT4 + s + 1) * 64 + lane];
			const int rel = lab - 32 * s - 4 * g;          // label's slot offset in this k-step
			float gv[8];
			#pragma unroll
			for (int i = 0; i < 4; ++i) {
				const bool la = rel == i, lb = rel == 16 + i;
				gv[i] = __builtin_fmaf(exp2_hw(__builtin_fmaf(pa[i], kLog2e, -mxl)), scale_g,
				                       la ? -inv_m : 0.f);
				gv[4 + i] = __builtin_fmaf(exp2_hw(__builtin_fmaf(pb[i], kLog2e, -mxl)), scale_g,
				                           lb ? -inv_m : 0.f);
				vlab += (la ? pa[i] : 0.f) + (lb ? pb[i] : 0.f);   // one match per lane at most
			}
			const v4u db = {bf16x2(gv[0], gv[1]), bf16x2(gv[2], gv[3]), bf16x2(gv[4], gv[5]),
			                bf16x2(gv[6], gv[7])};
			if (act) store_kstep<true>(out.dout, ld, ray0, 32 * s, kOutP, db, false, t, lane);
			#pragma unroll
			for (int kt = 0; kt < 4; ++kt) dh[kt] = mfma32(bu[kt], db, dh[kt]);
		};
		v4f pa0 = mfma32(w4s[(4 * s0 + 1) * 64 + lane], h3[0][1],
		                 mfma32(w4s[(4 * s0) * 64 + lane], h3[0][0], b4s[8 * s0 + g]));
		v4f pb0 = mfma32(w4s[(4 * s0 + 3) * 64 + lane], h3[0][1],
		                 mfma32(w4s[(4 * s0 + 2) * 64 + lane], h3[0][0], b4s[8 * s0 + 4 + g]));
		v4f pa1, pb1;
		Rec q0{w4s[(4 * s0 + 4) * 64 + lane], w4s[(4 * s0 + 5) * 64 + lane],
		       w4s[(4 * s0 + 6) * 64 + lane], w4s[(4 * s0 + 7) * 64 + lane], b4s[8 * s0 + 8 + g],
		       b4s[8 * s0 + 12 + g]}, q1;
		v4u bt0[4], bt1[4];
		#pragma unroll
		for (int kt = 0; kt < 4; ++kt) bt0[kt] = wqt[qW4 + (kt * kStepsT4 + s0) * 64 + lane];
		for (int s = s0; s < s0 + kClusters / 64; s += 2) {
			p2_step(s, pa0, pb0, pa1, pb1, q0, q1, bt0, bt1);
			p2_step(s + 1, pa1, pb1, pa0, pb0, q1, q0, bt1, bt0);
		}
		float lossc = 0.f;
		if (lab >= 0 && (lab >> 9) == half && g == ((lab >> 2) & 3)) lossc += log_z - vlab;
		if (half == 1) {
			// k-step 32: tile 64 (the log distance at feature 1024) and padding
			const bool dl = g == 0 && lab >= 0;
			const float e = dist - tgt;
			const v4u db = {bf16x2(dl ? 2.0f * e * inv_m : 0.f, 0.f), 0u, 0u, 0u};
			if (act) store_kstep<true>(out.dout, ld, ray0, 32 * (kClusters / 32), kOutP, db, false, t, lane);
			#pragma unroll
			for (int kt = 0; kt < 4; ++kt) dh[kt] = mfma32(bt0[kt], db, dh[kt]);
			if (dl) lossc += e * e;
			#pragma unroll
			for (int kt = 0; kt < 4; ++kt) xch[sl][kt][lane] = dh[kt];
		}
		#pragma unroll
		for (int off = kWave / 2; off > 0; off >>= 1) lossc += __shfl_xor(lossc, off);
		if (act && lane == 0) out.loss_part[2 * set + half] = lossc * inv_m;
		__syncthreads();
		if (!act) continue;
		if (half == 0) {
			// the hidden layers' data gradients
			#pragma unroll
			for (int kt = 0; kt < 4; ++kt) dh[kt] += xch[sl][kt][lane];
			v4u b3[8], b2[8];
			#pragma unroll
			for (int r = 0; r < 8; ++r) {
				b3[r] = wqt[qW3 + 64 * r + lane];
				b2[r] = wqt[qW2 + 64 * r + lane];
			}
			v4u dz[2];
			relu_back(dz, dh, h3[0]);
			store_hidden(out.dz3, ld, ray0, dz, t, lane);
			data_back(dh, dz, b3);
			relu_back(dz, dh, h2[0]);
			store_hidden(out.dz2, ld, ray0, dz, t, lane);
			data_back(dh, dz, b2);
			relu_back(dz, dh, h1[0]);
			store_hidden(out.dz1, ld, ray0, dz, t, lane);
		} else {
			// the layer inputs
			#pragma unroll
			for (int s = 0; s < 4; ++s)
				store_kstep(out.x, ld, ray0, 32 * s, kInP, xs[0][s], true, t, lane);
			store_hidden(out.h1, ld, ray0, h1[0], t, lane);
			store_hidden(out.h2, ld, ray0, h2[0], t, lane);
			store_hidden(out.h3, ld, ray0, h3[0], t, lane);
		}
	}
}

// dW and db of one layer for kGradNT row tiles of 16 (rows n0.., tiles at
// and beyond nt left at zero) and KT tiles of 16 inputs over the kSplitRays
// rays from r0: D[n][k] = sum_r dY[n][r] X[k][r], and db as the product
// with a ones column.  The k-steps are unrolled so that all operand loads
// can be in flight at once; the X operands of a k-step serve both row
// tiles.  acc[u][t]: row tile u, input tile t (t = KT: the bias column).
constexpr int kSplitRays = 256;             // rays per split of the gradient products
constexpr int kGradWaves = 4;               // splits (waves) per block, added in LDS

// dW and db of one layer for NT row tiles of 16 (rows n0.., tiles at and
// beyond nt read tile 0 and are not stored) and KT tiles of 16 inputs over
// the kSplitRays rays from r0: D[n][k] = sum_r dY[n][r] X[k][r], and db as
// the product with a ones column.  Every operand load is issued before the
// products (one memory latency for the whole split); the X operands of a
// k-step serve all NT row tiles.  acc[u][t]: row tile u, input tile t (t =
// KT: the bias column).
template <int KT, int NT>
__device__ inline void grad_tiles(v4f acc[NT][KT + 1], const uint16_t *__restrict__ dy,
                                  const uint16_t *__restrict__ x, int ld, int n0, int nt, int r0,
                                  int lane) {
	const int c = lane & 15;
	const unsigned one = c == 0 ? 0x3f803f80u : 0u;   // bf16 1.0 in column 0
	const v4u ones = {one, one, one, one};
	// operand records: 1 KiB each, lane l's 16 B at 16 l
	const size_t tstride = (size_t)(ld >> 5) * 512;               // one feature tile
	const uint16_t *arow[NT];
	#pragma unroll
	for (int u = 0; u < NT; ++u) arow[u] = dy + rec_offset(n0 + 16 * (u < nt ? u : 0), r0, ld) + 8 * lane;
	const uint16_t *brow = x + rec_offset(0, r0, ld) + 8 * lane;
	constexpr int kSteps = kSplitRays / 32;
	v4u a[kSteps][NT], b[kSteps][KT];
	#pragma unroll
	for (int q = 0; q < kSteps; ++q) {
		#pragma unroll
		for (int u = 0; u < NT; ++u) a[q][u] = *reinterpret_cast<const v4u *>(arow[u] + 512 * q);
		#pragma unroll
		for (int t = 0; t < KT; ++t)
			b[q][t] = *reinterpret_cast<const v4u *>(brow + t * tstride + 512 * q);
	}
	__builtin_amdgcn_sched_barrier(0);   // keep the scheduler from interleaving them again
	#pragma unroll
	for (int q = 0; q < kSteps; ++q)
		#pragma unroll
		for (int u = 0; u < NT; ++u) {
			#pragma unroll
			for (int t = 0; t < KT; ++t) acc[u][t] = mfma32(a[q][u], b[q][t], acc[u][t]);
			acc[u][KT] = mfma32(a[q][u], ones, acc[u][KT]);
		}
}

// One block's job: the waves' sums over their splits, added in LDS (red:
// [wave][entry][lane]) in wave order and stored as the block's partial
// gradient (wave w stores every kGradWaves-th entry from w).
template <int KT, int NT>
__device__ inline void grad_group(float *__restrict__ part, float *red, const uint16_t *dy,
                                  const uint16_t *x, int ld, int n0, int nt, int r0, int r1,
                                  int ow, int ldw, int ob, bool layer1, int wave, int lane) {
	constexpr int kEntries = NT * (KT + 1) * 4;
	v4f acc[NT][KT + 1];
	#pragma unroll
	for (int u = 0; u < NT; ++u)
		#pragma unroll
		for (int t = 0; t <= KT; ++t) acc[u][t] = (v4f){0.f, 0.f, 0.f, 0.f};
	for (int r = r0; r < r1; r += kSplitRays) grad_tiles<KT, NT>(acc, dy, x, ld, n0, nt, r, lane);
	#pragma unroll
	for (int u = 0; u < NT; ++u)
		#pragma unroll
		for (int t = 0; t <= KT; ++t)
			#pragma unroll
			for (int i = 0; i < 4; ++i)
				red[((size_t)wave * kEntries + (u * (KT + 1) + t) * 4 + i) * kWave + lane] = acc[u][t][i];
	__syncthreads();
	const int g = lane >> 4, c = lane & 15;
	for (int e = wave; e < kEntries; e += kGradWaves) {
		float v = red[e * kWave + lane];
		#pragma unroll
		for (int w = 1; w < kGradWaves; ++w) v += red[(w * kEntries + e) * kWave + lane];
		const int u = e / ((KT + 1) * 4), t = (e / 4) % (KT + 1), i = e % 4;
		if (u >= nt) continue;
		const int n = n0 + 16 * u + 4 * g + i;
		if (t < KT) {
			const int k = 16 * t + c;
			const int colw = layer1 ? (k < 96 ? k + 6 : (k < kIn ? k - 96 : k)) : k;
			part[ow + n * ldw + colw] = v;
		} else if (c == 0) {
			part[ob + n] = v;
		}
	}
}

// Block = (group of row tiles, kGradWaves consecutive wave ranges of
// kWaveRays rays); wave w takes range kGradWaves * blk + w (empty beyond the
// rows).  Groups: W4 in pairs of row tiles (0..32, the last one tile), then
// pairs of W3 (33, 34) and W2 (35, 36), then single row tiles of W1
// (37..40), whose 7 input tiles would not fit a pair in the register
// budget.  The block's partial goes to part[blk].
constexpr int kWaveRays = 2 * kSplitRays;   // 512 rays per wave: one block per CU at 8,192 rays
constexpr int kGroups4 = (kOutP / 16 + 1) / 2;                     // 33
constexpr int kGradGroups = kGroups4 + 2 + 2 + 4;                  // 41
constexpr int kRedFloats = kGradWaves * 2 * (kHid / 16 + 1) * 4 * kWave;  // 40 KiB
__global__ __launch_bounds__(kGradWaves * kWave) void grad_partials_kernel(
    float *__restrict__ part, TrainRows in, int splits) {
	__shared__ float red[kRedFloats];
	const int n_pad = splits * kSplitRays;
	const int blks = (n_pad + kGradWaves * kWaveRays - 1) / (kGradWaves * kWaveRays);
	const int grp = blockIdx.x / blks, blk = blockIdx.x % blks;
	const int wave = (int)(threadIdx.x / kWave), lane = __lane_id();
	const int rb = (kGradWaves * blk + wave) * kWaveRays;
	const int r0 = rb < n_pad ? rb : n_pad, r1 = rb + kWaveRays < n_pad ? rb + kWaveRays : n_pad;
	float *pp = part + (size_t)blk * kParams;
	if (grp < kGroups4) {
		const int t0 = 2 * grp, nt = kOutP / 16 - t0 < 2 ? kOutP / 16 - t0 : 2;
		grad_group<kHid / 16, 2>(pp, red, in.dout, in.h3, in.ld, 16 * t0, nt, r0, r1, oW4,
		                         kHid, oB4, false, wave, lane);
	} else if (grp < kGroups4 + 2) {
		grad_group<kHid / 16, 2>(pp, red, in.dz3, in.h2, in.ld, 32 * (grp - kGroups4), 2, r0,
		                         r1, oW3, kHid, oB3, false, wave, lane);
	} else if (grp < kGroups4 + 4) {
		grad_group<kHid / 16, 2>(pp, red, in.dz2, in.h1, in.ld, 32 * (grp - kGroups4 - 2), 2, r0,
		                         r1, oW2, kHid, oB2, false, wave, lane);
	} else {
		grad_group<kInP / 16, 1>(pp, red, in.dz1, in.x, in.ld, 16 * (grp - kGroups4 - 4), 1, r0,
		                         r1, oW1, kInP, oB1, true, wave, lane);
	}
}

// gradient = sum of the n_part partials (in order), Adam with the step word, the bf16
// copy and packs; block 0 also adds the sets' losses to *loss
__global__ __launch_bounds__(256) void train_adam_kernel(
    float *__restrict__ p, float *__restrict__ mo, float *__restrict__ ve,
    float *__restrict__ grads, uint16_t *__restrict__ wb, uint16_t *__restrict__ wp,
    uint16_t *__restrict__ wpt, const float *__restrict__ part, int n_part, float lr,
    const uint32_t *__restrict__ step_word, const float *__restrict__ loss_part, int n_sets,
    float *__restrict__ loss) {
	if (blockIdx.x == 0) {
		__shared__ float red[256];
		float sum = 0.f;
		for (int k = threadIdx.x; k < n_sets; k += 256) sum += loss_part[k];
		red[threadIdx.x] = sum;
		__syncthreads();
		for (int w = 128; w > 0; w >>= 1) {
			if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
			__syncthreads();
		}
		if (threadIdx.x == 0) *loss += red[0];
	}
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kParams) return;
	// the partials 8 at a time, all 8 loads in flight, added in order
	float gi = 0.f;
	for (int s0 = 0; s0 < n_part; s0 += 8) {
		float v[8];
		#pragma unroll
		for (int j = 0; j < 8; ++j) v[j] = s0 + j < n_part ? part[(size_t)(s0 + j) * kParams + i] : 0.f;
		#pragma unroll
		for (int j = 0; j < 8; ++j) gi += v[j];
	}
	grads[i] = gi;
	float c1, c2;
	bias_corrections(step_word, c1, c2);
	store_param(i, adam_step(p, mo, ve, i, gi, lr, c1, c2), wb, wp, wpt);
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
	uint16_t *wpack = nullptr;   // forward MFMA pack of the weights
	uint16_t *wpackt = nullptr;  // backward MFMA pack (W4, W3, W2 transposed)
	TrainRows rows{};            // the fused training step's feature-major rows
	float *part = nullptr;       // partial gradients, one per 2,048 rays
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
	// pack slots without a weight (W1's padding columns, W4^T beyond row
	// 1039) stay zero
	RTA_HIP_TRY(hipMemset(p->wpack, 0, (size_t)kPackRecords * 16));
	RTA_HIP_TRY(hipMemset(p->wpackt, 0, (size_t)kPackTRecords * 16));
	hipLaunchKernelGGL(store_params_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, 0, p->wb,
	                   p->wpack, p->wpackt, p->params);
	RTA_LAUNCH_CHECK("store_params_kernel");
	int dev = 0;
	RTA_HIP_TRY(hipGetDevice(&dev));
	RTA_HIP_TRY(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev));
	if (p->cus < 1) p->cus = 1;
	RTA_HIP_TRY(hipDeviceSynchronize());
	return RTA_OK;
}

inline float3 box_scale(const float *mn, const float *mx) {
	return make_float3(2.0f / (mx[0] - mn[0]), 2.0f / (mx[1] - mn[1]), 2.0f / (mx[2] - mn[2]));
}

// the three launches of one fused training step on rays [0, n)
int train_step(State *p, const float *o, const float *d, const int32_t *label,
               const float *target, int n, float3 lo, float3 sc, float lr, uint32_t *step,
               float *loss, hipStream_t st) {
	// rows padded to whole splits: the padding rays are rays without a hit
	// (zero output gradients), so every split is a full kSplitRays
	const int n_pad = (n + kSplitRays - 1) / kSplitRays * kSplitRays;
	const int n_sets = n_pad / 16;
	const long long need = (n_sets + kTrainSets - 1) / kTrainSets;
	const unsigned blocks = (unsigned)(need < p->cus ? need : p->cus);
	hipLaunchKernelGGL(train_rows_kernel<kTrainSets>, dim3(blocks), dim3(2 * kTrainSets * kWave), 0,
	                   st, reinterpret_cast<const v4u *>(p->wpack),
	                   reinterpret_cast<const v4u *>(p->wpackt), p->params, o, d, label, target, n,
	                   n_sets, lo, sc, 1.0f / (float)n, p->rows, step);
	RTA_LAUNCH_CHECK("train_rows_kernel");
	const int splits = n_pad / kSplitRays;
	const int blks = (n_pad + kGradWaves * kWaveRays - 1) / (kGradWaves * kWaveRays);
	hipLaunchKernelGGL(grad_partials_kernel, dim3(kGradGroups * blks), dim3(kGradWaves * kWave), 0,
	                   st, p->part, p->rows, splits);
	RTA_LAUNCH_CHECK("grad_partials_kernel");
	hipLaunchKernelGGL(train_adam_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, st,
	                   p->params, p->mom, p->vel, p->grads, p->wb, p->wpack, p->wpackt, p->part,
	                   blks, lr, step, p->rows.loss_part, 2 * n_sets, loss);
	RTA_LAUNCH_CHECK("train_adam_kernel");
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
	rc |= p->alloc(p->wpackt, (size_t)kPackTRecords * 8);
	const size_t ld = ((size_t)max_batch + kSplitRays - 1) / kSplitRays * kSplitRays;
	p->rows.ld = (int)ld;
	rc |= p->alloc(p->rows.x, ld * kInP);
	rc |= p->alloc(p->rows.h1, ld * kHid);
	rc |= p->alloc(p->rows.h2, ld * kHid);
	rc |= p->alloc(p->rows.h3, ld * kHid);
	rc |= p->alloc(p->rows.dout, ld * kOutP);
	rc |= p->alloc(p->rows.dz3, ld * kHid);
	rc |= p->alloc(p->rows.dz2, ld * kHid);
	rc |= p->alloc(p->rows.dz1, ld * kHid);
	rc |= p->alloc(p->rows.loss_part, ld / 8);
	rc |= p->alloc(p->part, (ld + kGradWaves * kWaveRays - 1) / (kGradWaves * kWaveRays) * kParams);
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
	hipLaunchKernelGGL(store_params_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, 0, p->wb,
	                   p->wpack, p->wpackt, p->params);
	RTA_LAUNCH_CHECK("store_params_kernel");
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
	                   p->params, p->mom, p->vel, p->grads, p->wb, p->wpack, p->wpackt, lr, c1, c2,
	                   nullptr);
	RTA_LAUNCH_CHECK("adam_kernel");
	return RTA_OK;
}

int rta_pred_adam_device(rta_pred h, float lr, uint32_t *step, rta_stream s) {
	State *p = reinterpret_cast<State *>(h);
	if (!p || !step) return invalid("rta_pred_adam_device: NULL argument");
	const int rc = rta_hip_add_u32(step, 1u, s);
	if (rc != RTA_OK) return rc;
	hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, as_stream(s),
	                   p->params, p->mom, p->vel, p->grads, p->wb, p->wpack, p->wpackt, lr, 0.0f,
	                   0.0f, step);
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

int rta_pred_train_step(rta_pred h, const float *o, const float *d, const int32_t *label,
                        const float *target, int n, const float *scene_min,
                        const float *scene_max, float lr, uint32_t *step, float *loss,
                        rta_stream s) {
	RTA_FAULT_POINT("rta_pred_train_step");
	State *p = reinterpret_cast<State *>(h);
	if (!p || !o || !d || !label || !target || !scene_min || !scene_max || !step || !loss)
		return invalid("rta_pred_train_step: NULL pointer");
	if (n <= 0 || n > p->max_batch) return invalid("rta_pred_train_step: n must be in 1..max_batch");
	return train_step(p, o, d, label, target, n, make_float3(scene_min[0], scene_min[1], scene_min[2]),
	                  box_scale(scene_min, scene_max), lr, step, loss, as_stream(s));
}

int rta_pred_train_batches(rta_pred h, const float *o, const float *d, const int32_t *label,
                           const float *target, int n, const int32_t *starts, int n_batches,
                           int batch, const float *scene_min, const float *scene_max, float lr,
                           uint32_t *step, float *loss, rta_stream s) {
	RTA_FAULT_POINT("rta_pred_train_batches");
	State *p = reinterpret_cast<State *>(h);
	if (!p || !o || !d || !label || !target || !scene_min || !scene_max || !step || !loss ||
	    (!starts && n_batches > 0))
		return invalid("rta_pred_train_batches: NULL pointer");
	if (batch <= 0 || batch > p->max_batch || n < 0 || n_batches < 0)
		return invalid("rta_pred_train_batches: batch must be in 1..max_batch, n and n_batches >= 0");
	for (int b = 0; b < n_batches; ++b)
		if (starts[b] < 0 || starts[b] >= n)
			return invalid("rta_pred_train_batches: a batch start outside [0, n)");
	const float3 lo = make_float3(scene_min[0], scene_min[1], scene_min[2]);
	const float3 sc = box_scale(scene_min, scene_max);
	hipStream_t st = as_stream(s);
	for (int b = 0; b < n_batches; ++b) {
		const size_t r = (size_t)starts[b];
		const int m = n - starts[b] < batch ? n - starts[b] : batch;
		const int rc = train_step(p, o + 3 * r, d + 3 * r, label + r, target + r, m, lo, sc, lr, step,
		                          loss, st);
		if (rc != RTA_OK) return rc;
	}
	return RTA_OK;
}

}  // extern "C"
