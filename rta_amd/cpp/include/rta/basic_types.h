// Core value types of the ray tracing interface (the roles of
// librta/basic_types.h:57-95 in the reference): a 104-byte triangle, an
// axis-aligned box, the 16-byte closest-hit record and the flat triangle
// list handed to plugins.  Byte layouts equal the C ABI's (rta_hip.h), so
// device buffers can be passed straight to the HIP entry points.
#pragma once

#include <float.h>
#include <stdint.h>

#include <string>

#include "../../../../include/rta_hip.h"

namespace rta {

typedef unsigned int uint;

struct vec2f {
	float x, y;
};
struct vec3f {
	float x, y, z;
};

template <typename tri_t_> struct basic_flat_triangle_list {
	typedef tri_t_ tri_t;
	uint triangles;
	tri_t *triangle;
	basic_flat_triangle_list() : triangles(0), triangle(0) {}
	explicit basic_flat_triangle_list(int size) : triangles(size), triangle(new tri_t[size]) {}
};

struct simple_triangle {
	typedef basic_flat_triangle_list<simple_triangle> input_flat_triangle_list_t;
	typedef vec3f vec3_t;
	vec3f a, b, c;
	vec3f na, nb, nc;
	vec2f ta, tb, tc;
	int material_index;
} __attribute__((aligned(8)));

struct simple_aabb {
	vec3f min, max;
};

template <typename tri_t_> struct triangle_intersection {
	typedef tri_t_ tri_t;
	float t, beta, gamma;
	uint ref;
	triangle_intersection() : t(FLT_MAX), beta(0), gamma(0), ref(0) {}
	explicit triangle_intersection(uint r) : t(FLT_MAX), beta(0), gamma(0), ref(r) {}
	bool valid() const { return t != FLT_MAX; }
	void reset() { t = FLT_MAX; ref = 0; }
	void barycentric_coord(vec3f *to) const {
		to->x = 1.0f - beta - gamma;
		to->y = beta;
		to->z = gamma;
	}
};

static_assert(sizeof(simple_triangle) == sizeof(rta_triangle), "triangle layout");
static_assert(sizeof(simple_aabb) == sizeof(rta_aabb), "box layout");
static_assert(sizeof(triangle_intersection<simple_triangle>) == sizeof(rta_hit), "hit layout");

}  // namespace rta
