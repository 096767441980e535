// Device self-test of the division-free slab arithmetic: div_cr against the
// GPU's correctly rounded IEEE division over seeded operands drawn from the
// admitted ranges (exact.h).  Run by the tests on every MI355X box.
#include "common.hpp"

namespace rta_hip {

__device__ inline float bits_to_float(uint32_t sign, int exp, uint32_t mant) {
	return __uint_as_float((sign << 31) | ((uint32_t)(exp + 127) << 23) | (mant & 0x7fffffu));
}

__global__ void check_division_kernel(uint64_t n, uint64_t seed, unsigned long long *bad,
                                      float *first_bad) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	unsigned long long local = 0;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint64_t h = mix64(seed ^ (2 * i)), g = mix64(seed ^ (2 * i + 1));
		float b, a;
		if ((i & 3) == 1) {         // whole admitted range
			b = bits_to_float(h & 1, (int)((h >> 1) % 89) - 60, (uint32_t)(h >> 8));
			a = bits_to_float(g & 1, (int)((g >> 1) % 157) - 96, (uint32_t)(g >> 8));
		} else if ((i & 3) == 3) {  // smallest quotients: tiny a over large b
			b = bits_to_float(h & 1, 28 - (int)((h >> 1) % 12), (uint32_t)(h >> 8));
			a = bits_to_float(g & 1, -96 + (int)((g >> 1) % 12), (uint32_t)(g >> 8));
		} else {                    // slab-like: coordinate differences over unit components
			b = u01(h) * 2.0f - 1.0f;
			a = (u01(g) * 4.0f - 2.0f) - (u01(mix64(g)) * 4.0f - 2.0f);
		}
		if (b == 0.0f || !divisor_fast_ok(b) || !dividend_fast_ok(a)) continue;
		const float y = 1.0f / b;
		const float q = div_cr(a, b, y);
		const float r = a / b;
		const bool same = __float_as_uint(q) == __float_as_uint(r) || (q != q && r != r) ||
		                  (q == 0.0f && r == 0.0f);
		if (!same) {
			++local;
			first_bad[0] = a;
			first_bad[1] = b;
		}
	}
	if (local) atomicAdd(bad, local);
}

// Enumeration behind div_cr1: thread x takes the divisor significand
// B = 1 + (mb_begin + x) * 2^-23 and runs over all 2^23 dividend significands
// A in [1, 2), comparing div_cr1(A, B, RN(1/B)) with the device's correctly
// rounded IEEE quotient.  Pure register arithmetic; one global atomic per
// thread at the end.
__global__ __launch_bounds__(256) void check_division_pairs_kernel(
    uint32_t mb_begin, uint32_t mb_count, unsigned long long *bad, uint32_t *example) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= mb_count) return;
	const uint32_t mb = mb_begin + x;
	const float b = __uint_as_float(0x3f800000u | mb);
	const float y = 1.0f / b;
	uint32_t local = 0, ex = 0;
	for (uint32_t ma = 0; ma < (1u << 23); ++ma) {
		const float a = __uint_as_float(0x3f800000u | ma);
		const bool differ = __float_as_uint(div_cr1(a, b, y)) != __float_as_uint(a / b);
		local += differ ? 1u : 0u;
		ex = differ ? ma : ex;
	}
	if (local) {
		atomicAdd(bad, (unsigned long long)local);
		example[0] = 0x3f800000u | ex;
		example[1] = 0x3f800000u | mb;
	}
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" int rta_hip_check_division_pairs(uint32_t mb_begin, uint32_t mb_count,
                                            uint64_t *mismatches, uint32_t *example) {
	if (!mismatches) return invalid("rta_hip_check_division_pairs: mismatches is NULL");
	if (mb_begin >= (1u << 23) || mb_count > (1u << 23) - mb_begin)
		return invalid("rta_hip_check_division_pairs: significand range beyond 2^23");
	*mismatches = 0;
	if (mb_count == 0) return RTA_OK;
	unsigned long long *d_bad = nullptr;
	uint32_t *d_ex = nullptr;
	RTA_HIP_TRY(hipMalloc(&d_bad, sizeof(unsigned long long)));
	hipError_t e = hipMalloc(&d_ex, 2 * sizeof(uint32_t));
	if (e != hipSuccess) {
		(void)hipFree(d_bad);
		return hip_status(e, "hipMalloc");
	}
	int rc = RTA_OK;
	e = hipMemset(d_bad, 0, sizeof(unsigned long long));
	if (e == hipSuccess) e = hipMemset(d_ex, 0, 2 * sizeof(uint32_t));
	if (e == hipSuccess) {
		hipLaunchKernelGGL(check_division_pairs_kernel, dim3(grid_for(mb_count, 256)), dim3(256), 0,
		                   0, mb_begin, mb_count, d_bad, d_ex);
		e = hipGetLastError();
	}
	unsigned long long bad = 0;
	uint32_t ex[2] = {0, 0};
	if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost);
	if (e == hipSuccess) e = hipMemcpy(ex, d_ex, sizeof ex, hipMemcpyDeviceToHost);
	if (e != hipSuccess) rc = hip_status(e, "rta_hip_check_division_pairs");
	*mismatches = bad;
	if (example) { example[0] = ex[0]; example[1] = ex[1]; }
	(void)hipFree(d_bad);
	(void)hipFree(d_ex);
	return rc;
}

extern "C" int rta_hip_check_division(uint64_t n, uint64_t seed, uint64_t *mismatches,
                                      float *example) {
	if (!mismatches) return invalid("rta_hip_check_division: mismatches is NULL");
	unsigned long long *d_bad = nullptr;
	float *d_ex = nullptr;
	RTA_HIP_TRY(hipMalloc(&d_bad, sizeof(unsigned long long)));
	hipError_t e = hipMalloc(&d_ex, 2 * sizeof(float));
	if (e != hipSuccess) {
		(void)hipFree(d_bad);
		return hip_status(e, "hipMalloc");
	}
	int rc = RTA_OK;
	e = hipMemset(d_bad, 0, sizeof(unsigned long long));
	if (e == hipSuccess) e = hipMemset(d_ex, 0, 2 * sizeof(float));
	if (e == hipSuccess) {
		hipLaunchKernelGGL(check_division_kernel, dim3(4096), dim3(256), 0, 0, n, seed, d_bad, d_ex);
		e = hipGetLastError();
	}
	unsigned long long bad = 0;
	float ex[2] = {0, 0};
	if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost);
	if (e == hipSuccess) e = hipMemcpy(ex, d_ex, sizeof ex, hipMemcpyDeviceToHost);
	if (e != hipSuccess) rc = hip_status(e, "rta_hip_check_division");
	*mismatches = bad;
	if (example) { example[0] = ex[0]; example[1] = ex[1]; }
	(void)hipFree(d_bad);
	(void)hipFree(d_ex);
	return rc;
}
