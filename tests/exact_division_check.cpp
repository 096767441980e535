// Bit-for-bit check of rta_hip::div_cr and rta_hip::div_cr1
// (rta_amd/csrc/hip/exact.h) against IEEE-754 binary32 division on the host.
// Built by tests/test_exact_division.py with g++ -O2 -mfma
// -ffp-contract=off, so both fmaf and '/' are single correctly rounded
// operations.
// Usage: exact_division_check <random_pairs> <seed>; prints the number of
// checked pairs and the mismatches of div_cr and of div_cr1, exit status 1
// on any mismatch.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../rta_amd/csrc/hip/exact.h"

using rta_hip::div_cr;
using rta_hip::div_cr1;
using rta_hip::dividend_fast_ok;
using rta_hip::divisor_fast_ok;

static uint64_t s_state;
static inline uint64_t next64() {
	uint64_t z = (s_state += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static inline float make_float(uint32_t sign, int exp, uint32_t mant) {
	uint32_t bits = (sign << 31) | ((uint32_t)(exp + 127) << 23) | (mant & 0x7fffff);
	float f;
	memcpy(&f, &bits, 4);
	return f;
}

static inline uint32_t bits_of(float f) {
	uint32_t u;
	memcpy(&u, &f, 4);
	return u;
}

static uint64_t checked = 0, bad = 0, bad1 = 0;

static inline bool same_quotient(float q, float r) {
	return bits_of(q) == bits_of(r) || (isnan(q) && isnan(r)) || (q == 0.0f && r == 0.0f);
}

static inline void check(float a, float b) {
	if (b == 0.0f || !divisor_fast_ok(b) || !dividend_fast_ok(a)) return;
	const float y = 1.0f / b;
	const float q = div_cr(a, b, y);
	const float q1 = div_cr1(a, b, y);
	const float r = a / b;
	++checked;
	if (!same_quotient(q, r)) {
		if (bad < 10) fprintf(stderr, "div_cr mismatch a=%a b=%a fast=%a ieee=%a\n", a, b, q, r);
		++bad;
	}
	if (!same_quotient(q1, r)) {
		if (bad1 < 10)
			fprintf(stderr, "div_cr1 mismatch a=%a b=%a fast=%a ieee=%a\n", a, b, q1, r);
		++bad1;
	}
}

int main(int argc, char **argv) {
	const uint64_t n = argc > 1 ? strtoull(argv[1], 0, 10) : 100000000ull;
	s_state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;

	// 1. random operands over the whole admitted range
	for (uint64_t i = 0; i < n; ++i) {
		uint64_t h = next64(), g = next64();
		float b = make_float(h & 1, (int)((h >> 1) % 89) - 60, (uint32_t)(h >> 8));
		float a = make_float(g & 1, (int)((g >> 1) % 157) - 96, (uint32_t)(g >> 8));
		check(a, b);
	}
	// 1b. the corner closest to the subnormal boundary: smallest dividends
	// over largest divisors (quotients 2^-124 .. 2^-100)
	for (uint64_t i = 0; i < n / 4; ++i) {
		uint64_t h = next64(), g = next64();
		float b = make_float(h & 1, 28 - (int)((h >> 1) % 12), (uint32_t)(h >> 8));
		float a = make_float(g & 1, -96 + (int)((g >> 1) % 12), (uint32_t)(g >> 8));
		check(a, b);
	}
	// 2. slab-test-like operands: coordinate differences over unit-vector
	// components
	for (uint64_t i = 0; i < n / 4; ++i) {
		uint64_t h = next64(), g = next64();
		float d = (float)((double)(h >> 11) / 9007199254740992.0 * 2.0 - 1.0);
		float lo = (float)((double)(g >> 11) / 9007199254740992.0 * 4.0 - 2.0);
		float o = (float)((double)(next64() >> 11) / 9007199254740992.0 * 4.0 - 2.0);
		check(lo - o, d);
	}
	// 3. quotients next to representable values: a = RN(b*q) +- k ulp
	for (uint64_t i = 0; i < n / 4; ++i) {
		uint64_t h = next64(), g = next64();
		float b = make_float(h & 1, (int)((h >> 1) % 41) - 20, (uint32_t)(h >> 8));
		float q = make_float(g & 1, (int)((g >> 1) % 41) - 20, (uint32_t)(g >> 8));
		float a = b * q;
		int k = (int)(next64() % 5) - 2;
		uint32_t ab = bits_of(a) + (uint32_t)k;
		memcpy(&a, &ab, 4);
		check(a, b);
	}
	// 4. every divisor mantissa in [1,2) against a fixed set of dividends
	const float as[] = {1.0f, 3.0f, 0.1f, 1.9999999f, 1.0000001f, 1e-20f, 7e15f};
	for (uint32_t m = 0; m < (1u << 23); ++m) {
		float b = make_float(0, 0, m);
		for (float a : as) check(a, b);
	}
	// 5. the exponent-scaling argument behind div_cr1's enumeration: one
	// significand pair (A, B) at every admitted exponent pair, at the
	// extremes of the ranges included (|b| = 2^-60 .. 2^28, |a| = 2^-96 ..
	// 2^60, so quotients from 2^-124 to 2^120 and residuals down to the
	// subnormal range)
	for (uint64_t i = 0; i < n / 64 + 1; ++i) {
		const uint32_t ma = (uint32_t)next64(), mb = (uint32_t)next64();
		for (int eb = -60; eb <= 27; eb += 1 + (int)(next64() % 8))
			for (int ea = -96; ea <= 59; ea += 1 + (int)(next64() % 16))
				check(make_float(0, ea, ma), make_float(0, eb, mb));
		check(make_float(0, -96, ma), make_float(0, 27, mb));
		check(make_float(0, 59, ma), make_float(0, -60, mb));
	}
	printf("checked %llu mismatches %llu mismatches1 %llu\n", (unsigned long long)checked,
	       (unsigned long long)bad, (unsigned long long)bad1);
	return (bad || bad1) ? 1 : 0;
}
