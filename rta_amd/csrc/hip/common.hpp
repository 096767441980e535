// Shared device helpers for the bbvh-hip kernels (gfx950, wave64).
//
// Every translation unit of this library is compiled with -ffp-contract=off
// and HIP's correctly rounded fp32 divide/sqrt, so each +,-,*,/ below is one
// IEEE-754 binary32 operation with round-to-nearest-even, exactly as the
// reference's host code computes it (librta/intersect.h, gcc SSE math).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <string>

#include "../../../include/rta_hip.h"
#include "exact.h"

namespace rta_hip {

static_assert(sizeof(rta_triangle) == 104, "triangle layout");
static_assert(sizeof(rta_aabb) == 24, "aabb layout");
static_assert(sizeof(rta_node) == 32, "node layout");
static_assert(sizeof(rta_hit) == 16, "hit layout");
static_assert(sizeof(rta_node_pair) == 64, "pair layout");
static_assert(sizeof(rta_treelet) == 32, "treelet layout");

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors

void set_error(const std::string &msg);
int hip_status(hipError_t e, const char *what);

#define RTA_HIP_TRY(call)                                             \
	do {                                                              \
		hipError_t e__ = (call);                                      \
		if (e__ != hipSuccess) return ::rta_hip::hip_status(e__, #call); \
	} while (0)

#define RTA_LAUNCH_CHECK(name)                                        \
	do {                                                              \
		hipError_t e__ = hipGetLastError();                           \
		if (e__ != hipSuccess) return ::rta_hip::hip_status(e__, name); \
	} while (0)

int invalid(const char *msg);

// a named profiler range (rta_hip_range_push/pop) for the enclosing scope
struct ScopedRange {
	explicit ScopedRange(const char *name) { rta_hip_range_push(name); }
	~ScopedRange() { rta_hip_range_pop(); }
	ScopedRange(const ScopedRange &) = delete;
	ScopedRange &operator=(const ScopedRange &) = delete;
};

// true (and *status set) when a test armed a fault for this entry point
bool take_fault(const char *entry, int *status);
#define RTA_FAULT_POINT(entry)                                        \
	do {                                                              \
		int st__ = 0;                                                 \
		if (::rta_hip::take_fault(entry, &st__)) return st__;         \
	} while (0)

inline hipStream_t as_stream(rta_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(long long n, int block) {
	return (unsigned)((n + block - 1) / block);
}

// --------------------------------------------------- exact fp helpers

// min/max exactly as `(a<b)?a:b` / `(a>b)?a:b` (the reference's MIN/MAX
// macros): for equal values the second operand wins, NaN selects b.
__device__ __host__ inline float sel_min(float a, float b) { return (a < b) ? a : b; }
__device__ __host__ inline float sel_max(float a, float b) { return (a > b) ? a : b; }

// float -> uint32 as CUDA's cvt.rzi.u32.f32 (the reference's implicit
// conversion `uint x = center.x * d.x`): truncate toward zero, saturate,
// NaN -> 0.
__device__ __host__ inline uint32_t f2u_rz_sat(float f) {
	if (!(f > 0.0f)) return 0u;            // NaN, -0, +0, negatives
	if (f >= 4294967296.0f) return 0xffffffffu;
	return (uint32_t)f;
}

__device__ __host__ inline uint32_t morton_encode_10bit(uint32_t x, uint32_t y, uint32_t z) {
	// interleave the low 10 bits of x, y, z; x in bit 0
	// (Morton_3D_Encode_10bit, bbvh-cuda/lbvh-kernels.cu:99-129)
	auto spread = [](uint32_t v) {
		v &= 0x3ffu;
		v = (v | (v << 16)) & 0x030000ffu;
		v = (v | (v << 8)) & 0x0300f00fu;
		v = (v | (v << 4)) & 0x030c30c3u;
		v = (v | (v << 2)) & 0x09249249u;
		return v;
	};
	return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

// Slab box test, librta/intersect.h:56-150, in branch-free form.  Once an
// axis rejects the result is false whatever the later axes compute, and
// when no axis rejects the sequence of t_near/t_far updates is the
// reference's, so (hit, dist) are identical.
__device__ inline bool box_test(float lo_x, float lo_y, float lo_z,
                                float hi_x, float hi_y, float hi_z,
                                float ox, float oy, float oz,
                                float dx, float dy, float dz, float &dist) {
	float tn = -FLT_MAX, tf = FLT_MAX;
	bool ok = true;
#define RTA_SLAB(LO, HI, O, D)                                  \
	{                                                           \
		if (D == 0.0f) {                                        \
			if (O < LO || O > HI) ok = false;                   \
		} else {                                                \
			float t1 = (LO - O) / D;                            \
			float t2 = (HI - O) / D;                            \
			if (t1 > t2) { float tmp = t1; t1 = t2; t2 = tmp; } \
			if (t1 > tn) tn = t1;                               \
			if (t2 < tf) tf = t2;                               \
			if (tn > tf) ok = false;                            \
			if (tf < 0.0f) ok = false;                          \
		}                                                       \
	}
	RTA_SLAB(lo_x, hi_x, ox, dx)
	RTA_SLAB(lo_y, hi_y, oy, dy)
	RTA_SLAB(lo_z, hi_z, oz, dz)
#undef RTA_SLAB
	dist = tn;
	return ok;
}

// The slab test for a ray none of whose direction components is zero, with
// every quotient formed by div_cr1 from the ray's reciprocal direction
// (yx = RN(1/dx), ...), without branches.  With no zero component all three
// axes take the reference's division branch, and:
//  - `if (t1 > t2) swap` leaves (t1, t2) = (min, max) of the two quotients
//    (equal values are interchangeable), so near = fminf, far = fmaxf;
//  - `if (t1 > t_near) t_near = t1` over the axes, from -FLT_MAX, is the
//    maximum of the finite near values (fmaxf may pick the other zero of
//    -0/+0, which no later comparison can tell apart), likewise t_far;
//  - t_near only grows and t_far only shrinks, so some axis has
//    t_near > t_far or t_far < 0 after its update exactly when the final
//    values do.
// Hence (hit, dist) equal box_test's whenever every quotient is exact:
// coord_fast_ok boxes and origins, ray_nz_ok directions.
__device__ inline bool box_test_nz(float lo_x, float lo_y, float lo_z,
                                   float hi_x, float hi_y, float hi_z,
                                   float ox, float oy, float oz,
                                   float dx, float dy, float dz,
                                   float yx, float yy, float yz, float &dist) {
	const float ax = div_cr1(lo_x - ox, dx, yx), bx = div_cr1(hi_x - ox, dx, yx);
	const float ay = div_cr1(lo_y - oy, dy, yy), by = div_cr1(hi_y - oy, dy, yy);
	const float az = div_cr1(lo_z - oz, dz, yz), bz = div_cr1(hi_z - oz, dz, yz);
	const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
	const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
	dist = tn;
	return tn <= tf && tf >= 0.0f;
}

// A coordinate is "fast-safe" if it is 0 or 2^-73 <= |x| <= 2^59.  For two
// such values lo - o is 0 or a multiple of 2^-96 of magnitude <= 2^60, so
// every slab numerator is admitted by dividend_fast_ok.
__device__ __host__ inline bool coord_fast_ok(float x) {
	const float fx = fabsf(x);
	return x == 0.0f || (fx >= 0x1p-73f && fx <= 0x1p59f);
}

// A ray may use box_test_nz (given fast-safe boxes) when its origin is
// fast-safe and every direction component is nonzero with magnitude in
// [2^-60, 2^28].  Other rays (axis-parallel ones among them) take the IEEE
// box_test.
__device__ inline bool ray_nz_ok(float ox, float oy, float oz, float dx, float dy, float dz) {
	return coord_fast_ok(ox) && coord_fast_ok(oy) && coord_fast_ok(oz) && dx != 0.0f &&
	       dy != 0.0f && dz != 0.0f && divisor_fast_ok(dx) && divisor_fast_ok(dy) &&
	       divisor_fast_ok(dz);
}

// intersect_tri_opt, librta/intersect.h:8-53, operation for operation.
// (ea_x, ea_y, ea_z) = a - b and (ec_x, ...) = a - c per component, which
// is what the reference forms first (its `a,b,c` and `d,e,f`).
__device__ inline bool tri_test(float ax, float ay, float az,
                                float a, float b, float c,
                                float d, float e, float f,
                                float ox, float oy, float oz,
                                float g, float h, float i,
                                float &t_out, float &beta_out, float &gamma_out) {
	const float j = ax - ox;
	const float k = ay - oy;
	const float l = az - oz;

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

	// A quotient num/M can only be > 0 when num is nonzero and has M's sign
	// (a zero or underflowed quotient is not > 0 either).  Rejecting here
	// skips three divisions without changing any outcome; NaNs and M == 0
	// pass on to the divisions, which reject them as the reference does.
	const uint32_t sm = __float_as_uint(M);
	if (beta == 0.0f || gamma == 0.0f || tt == 0.0f ||
	    ((sm ^ __float_as_uint(beta)) | (sm ^ __float_as_uint(gamma)) | (sm ^ __float_as_uint(tt))) >> 31)
		return false;

	beta /= M;
	gamma /= M;
	tt /= M;

	if (tt > 0.0f && beta > 0.0f && gamma > 0.0f && beta + gamma <= 1.0f) {
		t_out = tt;
		beta_out = beta;
		gamma_out = gamma;
		return true;
	}
	return false;
}

// The bound on child-box entry distances that a seed imposes: a seed s is
// the t of a real hit of the ray, so the closest hit lies at t <= s, but a
// box holding that closest triangle can have a computed entry distance a
// little beyond its computed t (the slab quotients and Cramer's rule round
// differently; on planar triangles met at grazing angles the excess
// reaches about 2^-16.3 relative).  Pruning at s itself would then lose the
// closest hit; pruning beyond s + s * 2^-10 keeps a margin of 2^-10
// relative.  s + s * 2^-10 >= s for every s > 0, and FLT_MAX widens to
// +inf (no bound).
__device__ __host__ inline float seed_bound(float s) { return s + s * 0x1p-10f; }

// bbvh_node link word accessors (bbvh/bbvh-node.h:18-31)
__device__ __host__ inline bool link_inner(uint32_t w) { return (w & 1u) == 1u; }
__device__ __host__ inline uint32_t link_value(uint32_t w) { return w >> 1; }

// Counter-based RNG (splitmix64 finaliser over a 64-bit counter).
__device__ __host__ inline uint64_t mix64(uint64_t z) {
	z += 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}
// uniform float in [0,1) with 24 random bits: exact multiple of 2^-24
__device__ __host__ inline float u01(uint64_t bits) {
	return (float)(uint32_t)(bits >> 40) * (1.0f / 16777216.0f);
}

}  // namespace rta_hip
