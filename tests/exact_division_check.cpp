// Bit-for-bit check of rta_hip::div_cr (rta_amd/csrc/hip/exact.h) against
// IEEE-754 binary32 division on the host.  Built by
// tests/test_exact_division.py with g++ -O2 -mfma -ffp-contract=off, so both
// fmaf and '/' are single correctly rounded operations.
// Usage: exact_division_check <random_pairs> <seed>; prints the number of
// checked pairs and mismatches, exit status 1 on any mismatch.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../rta_amd/csrc/hip/exact.h"

using rta_hip::div_cr;
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

static uint64_t checked = 0, bad = 0;

static inline void check(float a, float b) {
	if (b == 0.0f || !divisor_fast_ok(b) || !dividend_fast_ok(a)) return;
	const float y = 1.0f / b;
	const float q = div_cr(a, b, y);
	const float r = a / b;
	++checked;
	bool same = bits_of(q) == bits_of(r) || (isnan(q) && isnan(r)) || (q == 0.0f && r == 0.0f);
	if (!same) {
		if (bad < 10)
			fprintf(stderr, "mismatch a=%a b=%a fast=%a ieee=%a\n", a, b, q, r);
		++bad;
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
	printf("checked %llu mismatches %llu\n", (unsigned long long)checked,
	       (unsigned long long)bad);
	return bad ? 1 : 0;
}
