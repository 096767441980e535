// The ray tracing interfaces a plugin implements and a host drives (the
// roles of librta/raytrav.h:43-437 in the reference): acceleration
// structure, its constructor, bouncer, ray generator and ray tracer, with
// basic_raytracer::trace() running the generate / (prepare, trace, bounce)*
// loop whose trace_rays() timings the host reports.
//
// Objects cross the plugin boundary by pointer, so every class here keeps
// the reference's virtual functions in the reference's declaration order
// (the same vtable slots) and the reference's data members in its order
// (the same offsets), including the reference's guard virtuals
// (please_derive_from_*, dont_forget_to_initialize_max_t) and raytracer's
// lack of a virtual destructor; tests/abi_layout_check.cpp checks both.
#pragma once

#include <math.h>

#include <algorithm>
#include <string>
#include <vector>

#include "basic_types.h"

namespace rta {

//! w x h pixels of C values of type T, row-major, channels interleaved (the
//! layout of the reference's image<T, C>, librta/image.h:15-31; no PNG I/O).
template <typename T, unsigned int C> class image {
public:
	image(unsigned int w, unsigned int h) : w(w), h(h), data(0) {
		if (w > 0 && h > 0) data = new T[(size_t)w * h * C];
	}
	image(const image &o) : w(o.w), h(o.h), data(0) {
		if (o.data) {
			data = new T[(size_t)w * h * C];
			std::copy(o.data, o.data + (size_t)w * h * C, data);
		}
	}
	image &operator=(const image &) = delete;
	~image() {
		delete[] data;
		w = h = 0;
	}
	unsigned int w, h;
	T *data;  //!< 0: no pixels
	T &pixel(unsigned int x, unsigned int y) { return data[((size_t)y * w + x) * C]; }
	T &pixel(unsigned int x, unsigned int y, unsigned int c) {
		return data[((size_t)y * w + x) * C + c];
	}
	const T &pixel(unsigned int x, unsigned int y) const { return data[((size_t)y * w + x) * C]; }
	const T &pixel(unsigned int x, unsigned int y, unsigned int c) const {
		return data[((size_t)y * w + x) * C + c];
	}
};

class acceleration_structure {
public:
	virtual ~acceleration_structure() {}
	virtual std::string identification() = 0;
	//! guard of the reference (raytrav.h:48): implemented by
	//! basic_acceleration_structure only
	virtual void please_derive_from_basic_acceleration_structure() = 0;
	virtual void dump_acceleration_structure(const std::string &) {}
	virtual void dump_primitives(const std::string &) {}
};

template <typename box_t_, typename tri_t_>
class basic_acceleration_structure : public acceleration_structure {
public:
	typedef box_t_ box_t;
	typedef tri_t_ tri_t;
	//! array indexable with triangle_intersection::ref (device memory for GPU structures)
	virtual tri_t *triangle_ptr() = 0;
	virtual int triangle_count() = 0;
	//! host-accessible copy of triangle_ptr(); may allocate (then release with
	//! free_canonical_triangles)
	virtual tri_t *canonical_triangle_ptr() { return triangle_ptr(); }
	virtual void free_canonical_triangles(tri_t *) {}
	void please_derive_from_basic_acceleration_structure() override {}
};

class acceleration_structure_constructor {
public:
	virtual ~acceleration_structure_constructor() {}
	virtual std::string identification() = 0;
	//! guard of the reference (raytrav.h:89)
	virtual void please_derive_from_basic_acceleration_structure_constructor() = 0;
};

template <typename box_t_, typename tri_t_>
class basic_acceleration_structure_constructor : public acceleration_structure_constructor {
public:
	typedef box_t_ box_t;
	typedef tri_t_ tri_t;
	typedef typename tri_t::input_flat_triangle_list_t flat_triangle_list_t;
	virtual basic_acceleration_structure<box_t, tri_t> *build(flat_triangle_list_t *tris) = 0;
	//! whether build() expects triangles in host memory
	virtual bool expects_host_triangles() { return true; }
	void please_derive_from_basic_acceleration_structure_constructor() override {}
};

class bouncer {
public:
	virtual ~bouncer() {}
	virtual void bounce() = 0;
	virtual bool trace_further_bounces() = 0;
	virtual void new_pass() {}
	virtual std::string identification() = 0;
};

//! Host ray storage: per pixel its origin and direction (interleaved, two
//! vec3f per pixel) and its max_t, as the reference stores them
//! (raytrav.h:270-298).
class ray_generator {
public:
	image<vec3f, 2> raydata;
	image<float, 1> ray_max_t;
	ray_generator(uint res_x, uint res_y) : raydata(res_x, res_y), ray_max_t(res_x, res_y) {}
	virtual ~ray_generator() {}
	virtual void generate_rays() = 0;
	uint res_x() const { return raydata.w; }
	uint res_y() const { return raydata.h; }
	vec3f *origin(uint x, uint y) { return &raydata.pixel(x, y, 0); }
	const vec3f *origin(uint x, uint y) const { return &raydata.pixel(x, y, 0); }
	vec3f *direction(uint x, uint y) { return &raydata.pixel(x, y, 1); }
	const vec3f *direction(uint x, uint y) const { return &raydata.pixel(x, y, 1); }
	float &max_t(uint x, uint y) { return ray_max_t.pixel(x, y, 0); }
	const float &max_t(uint x, uint y) const { return ray_max_t.pixel(x, y, 0); }
	virtual std::string identification() = 0;
	//! text dump of the rays (the .ray format, librta/raytrav.h:353-366)
	virtual void dump_rays(const std::string &) {}
	//! guard of the reference (raytrav.h:297)
	virtual void dont_forget_to_initialize_max_t() = 0;
};

//! Pinhole camera after Shirley (librta/raytrav.h:300-351).
class cam_ray_generator_shirley : public ray_generator {
protected:
	float aspect, fovy;
	vec3f position, dir, up;

public:
	cam_ray_generator_shirley(uint res_x, uint res_y)
	    : ray_generator(res_x, res_y), aspect(float(res_x) / float(res_y)), fovy(45.0f),
	      position{0, 0, 0}, dir{0, 0, -1}, up{0, 1, 0} {}
	void setup(const vec3f *new_position, const vec3f *new_dir, const vec3f *new_up,
	           float new_fovy) {
		position = *new_position;
		dir = *new_dir;
		up = *new_up;
		fovy = new_fovy;
	}
	void generate_rays() override;  // host version, in librta
	std::string identification() override { return "pinhole ray generator (Shirley)"; }
	void dont_forget_to_initialize_max_t() override {}
	float fov() const { return fovy; }
	float aspect_ratio() const { return aspect; }
	const vec3f &camera_position() const { return position; }
	const vec3f &camera_direction() const { return dir; }
	const vec3f &camera_up() const { return up; }
};

//! No virtual destructor, as in the reference (raytrav.h:373-383): tracers
//! are never deleted through this base, and a virtual destructor would add
//! two vtable slots ahead of trace().
class raytracer {
public:
	virtual void trace() = 0;
	virtual std::string identification() = 0;
	virtual raytracer *copy() = 0;
	//! whether the tracer honours the ray generator's max_t
	virtual bool supports_max_t() = 0;

protected:
	~raytracer() {}
};

template <typename box_t_, typename tri_t_> class basic_raytracer : public raytracer {
public:
	typedef box_t_ box_t;
	typedef tri_t_ tri_t;

protected:
	rta::ray_generator *raygen;
	rta::bouncer *bouncer;
	basic_acceleration_structure<box_t, tri_t> *accel_struct;
	virtual float trace_rays() = 0;

public:
	std::vector<float> timings;  //!< trace_rays() ms of each bounce of the last trace()
	basic_raytracer(rta::ray_generator *rg, rta::bouncer *b, basic_acceleration_structure<box_t, tri_t> *as)
	    : raygen(rg), bouncer(b), accel_struct(as) {}
	//! untimed preparation before each trace_rays (e.g. resetting hits)
	virtual void prepare_trace() {}
	void trace() override {
		raygen->generate_rays();
		bouncer->new_pass();
		timings.clear();
		do {
			prepare_trace();
			float ms = trace_rays();
			timings.push_back(ms);
			bouncer->bounce();
		} while (bouncer->trace_further_bounces());
	}
	virtual void ray_generator(rta::ray_generator *rg) { raygen = rg; }
	virtual void ray_bouncer(rta::bouncer *b) { bouncer = b; }
	virtual void acceleration_structure(basic_acceleration_structure<box_t, tri_t> *as) {
		accel_struct = as;
	}
	basic_acceleration_structure<box_t, tri_t> *acceleration_structure() { return accel_struct; }
	basic_raytracer *copy() override = 0;
};

}  // namespace rta
