// Exact fp32 arithmetic helpers shared by host and device code.
//
// div_cr(a, b, y): the correctly rounded quotient RN(a/b) computed from
// y = RN(1/b) with one multiply and four FMAs:
//   q0 = RN(a*y)                  (within 2 ulp of a/b)
//   q1 = q0 + (a - b*q0)*y        (faithful: within 1 ulp)
//   q2 = q1 + (a - b*q1)*y        (= RN(a/b))
// The last step is Markstein's theorem (Markstein 1990; Cornea, Harrison and
// Tang, "Scientific Computing on Itanium-based Systems", 2002): if y is
// within half an ulp of 1/b and q1 is within one ulp of a/b, the residual
// a - b*q1 is exact in an FMA and RN(q1 + r*y) = RN(a/b), provided nothing
// overflows or underflows.  divisor_fast_ok()/dividend_fast_ok() confine
// operands to a range where that holds: b == 0 is never divided (the slab
// test handles it), nonzero |b| in [2^-60, 2^28], |a| in [2^-96, 2^60] or
// a == 0.  Then every nonzero quotient lies in [2^-124, 2^120] and every
// rounded intermediate (q0, r0, q1, r1, q2) is a normal number; a quotient
// in the subnormal range would void the theorem (the fixed 2^-149 grid
// makes RN(q1 + r1*y) differ from RN(a/b) by one ulp for some operands).
// For a == 0 the result is a zero whose sign may differ from the IEEE
// quotient's; callers only compare it, where -0 == +0.
// The helper is checked bit-for-bit against IEEE division on the CPU and on
// the GPU (tests/test_exact_division.py).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define RTA_HD __host__ __device__
#else
#define RTA_HD
#endif

namespace rta_hip {

RTA_HD inline float div_cr(float a, float b, float y) {
	const float q0 = a * y;
	const float r0 = fmaf(-b, q0, a);
	const float q1 = fmaf(r0, y, q0);
	const float r1 = fmaf(-b, q1, a);
	return fmaf(r1, y, q1);
}

// div_cr1(a, b, y): RN(a/b) with one correction step (one multiply, two
// FMAs), for the same operand ranges as div_cr:
//   q0 = RN(a*y),  r0 = RN(a - b*q0),  q1 = RN(q0 + r0*y).
// No theorem covers this shorter sequence (q0 is not always faithful), so it
// rests on enumeration: for every pair of significands A, B in [1, 2) (2^46
// pairs) q1 == RN(A/B) (rta_hip_check_division_pairs, all 64 chunks).  The
// enumeration extends to every exponent pair in range because each step
// scales exactly by powers of two there: y, q0 and q1 stay normal (quotients
// in [2^-124, 2^120]), and r0 is either exact (the exact a - b*q0 is a
// multiple of 2^(e_a - 47) >= 2^-143, so it is representable even when
// subnormal) or it needs more than 24 bits, in which case
// |a - b*q0| >= 2^(e_a - 23) >= 2^-119 is normal and rounds as at unit scale.
RTA_HD inline float div_cr1(float a, float b, float y) {
	const float q0 = a * y;
	const float r0 = fmaf(-b, q0, a);
	return fmaf(r0, y, q0);
}

// |b| of a divisor usable with div_cr (zero is handled by the caller)
RTA_HD inline bool divisor_fast_ok(float b) {
	const float fb = fabsf(b);
	return b == 0.0f || (fb >= 0x1p-60f && fb <= 0x1p28f);
}

// numerator usable with div_cr; NaN passes (both paths give NaN)
RTA_HD inline bool dividend_fast_ok(float a) {
	const float fa = fabsf(a);
	return !((fa < 0x1p-96f && a != 0.0f) || fa > 0x1p60f);
}

}  // namespace rta_hip
