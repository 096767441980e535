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
