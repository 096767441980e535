// GPU glue between the interfaces and the HIP C ABI (the roles of
// librta/cuda.h:90-304 in the reference): device ray buffers and the GPU
// pinhole generator, device arrays, the device intersection buffer owned by
// GPU bouncers, and gpu_raytracer whose untimed prepare_trace() resets the
// intersections.  Every device operation goes through include/rta_hip.h.
#pragma once

#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../../include/rta_hip.h"
#include "raytrav.h"

namespace rta {
namespace hip {

//! Throws with the library's message when a C-ABI call fails (the
//! reference prints and exits, librta/cuda.h:70-78).
inline void check(int status, const char *what) {
	if (status != RTA_OK)
		throw std::runtime_error(std::string(what) + ": " + rta_hip_last_error());
}

//! Set by a GPU plugin's initialize(): the host then gives the tracer a GPU
//! bouncer (the reference's plugin_with_cuda_loaded, librta/cuda.cpp:10-13).
extern bool plugin_with_hip_loaded;
bool using_hip();

template <typename T> class device_array {
public:
	T *data;
	uint n;
	bool owns_memory;
	device_array() : data(0), n(0), owns_memory(true) {}
	device_array(const device_array &) = delete;
	device_array &operator=(const device_array &) = delete;
	void allocate(uint elems) {
		release();
		n = elems;
		void *p = 0;
		check(rta_hip_malloc(&p, sizeof(T) * (size_t)n), "rta_hip_malloc");
		data = (T *)p;
		owns_memory = true;
	}
	void upload(const T *in, uint elems) {
		allocate(elems);
		check(rta_hip_memcpy_htod(data, in, sizeof(T) * (size_t)n), "rta_hip_memcpy_htod");
	}
	T *download() const {
		T *t = new T[n];
		check(rta_hip_memcpy_dtoh(t, data, sizeof(T) * (size_t)n), "rta_hip_memcpy_dtoh");
		return t;
	}
	void release() {
		if (owns_memory && data) rta_hip_free(data);
		data = 0;
		n = 0;
	}
	~device_array() { release(); }
};

//! Device ray storage: origin[3*w*h], direction[3*w*h], max_t[w*h].
class gpu_ray_generator {
public:
	float *gpu_origin, *gpu_direction, *gpu_maxt;
	uint w, h;
	gpu_ray_generator(uint w, uint h) : gpu_origin(0), gpu_direction(0), gpu_maxt(0), w(w), h(h) {
		void *p = 0;
		check(rta_hip_malloc(&p, (size_t)w * h * 3 * sizeof(float)), "ray origins");
		gpu_origin = (float *)p;
		check(rta_hip_malloc(&p, (size_t)w * h * 3 * sizeof(float)), "ray directions");
		gpu_direction = (float *)p;
		check(rta_hip_malloc(&p, (size_t)w * h * sizeof(float)), "ray max_t");
		gpu_maxt = (float *)p;
	}
	virtual ~gpu_ray_generator() {
		rta_hip_free(gpu_origin);
		rta_hip_free(gpu_direction);
		rta_hip_free(gpu_maxt);
	}
};

//! The pinhole generator on the GPU (same rays as the host version).
class cam_ray_generator_shirley : public rta::cam_ray_generator_shirley, public gpu_ray_generator {
public:
	cam_ray_generator_shirley(uint res_x, uint res_y)
	    : rta::cam_ray_generator_shirley(res_x, res_y), gpu_ray_generator(res_x, res_y) {}
	void generate_rays() override {
		check(rta_setup_shirley_rays(gpu_direction, gpu_origin, gpu_maxt, fovy, aspect, w, h,
		                             &dir.x, &position.x, &up.x, FLT_MAX, 0),
		      "rta_setup_shirley_rays");
	}
	//! the reference's .ray text format: "ray", "w h", then per ray
	//! "id 2 ox oy oz dx dy dz 0 max_t" (librta/raytrav.h:353-366)
	void dump_rays(const std::string &filename) override {
		const size_t n = (size_t)w * h;
		std::vector<float> o(3 * n), d(3 * n), t(n);
		check(rta_hip_memcpy_dtoh(o.data(), gpu_origin, 12 * n), "download origins");
		check(rta_hip_memcpy_dtoh(d.data(), gpu_direction, 12 * n), "download directions");
		check(rta_hip_memcpy_dtoh(t.data(), gpu_maxt, 4 * n), "download max_t");
		std::ofstream out(filename.c_str());
		out << "ray\n" << w << " " << h << "\n";
		for (size_t i = 0; i < n; ++i)
			out << i << " " << 2 << " " << o[3 * i] << " " << o[3 * i + 1] << " " << o[3 * i + 2]
			    << " " << d[3 * i] << " " << d[3 * i + 1] << " " << d[3 * i + 2] << " 0 " << t[i]
			    << "\n";
	}
	std::string identification() override { return "HIP pinhole ray generator (Shirley)"; }
};

//! Rays uploaded from host arrays (e.g. a recorded bounce set).
class uploaded_rays : public rta::ray_generator, public gpu_ray_generator {
public:
	uploaded_rays(uint w, uint h) : rta::ray_generator(w, h), gpu_ray_generator(w, h) {}
	void generate_rays() override {
		const size_t n = (size_t)w * h;
		std::vector<vec3f> o(n), d(n);
		for (size_t i = 0; i < n; ++i) {
			o[i] = raydata.data[2 * i];
			d[i] = raydata.data[2 * i + 1];
		}
		check(rta_hip_memcpy_htod(gpu_origin, o.data(), 12 * n), "upload origins");
		check(rta_hip_memcpy_htod(gpu_direction, d.data(), 12 * n), "upload dirs");
		check(rta_hip_memcpy_htod(gpu_maxt, ray_max_t.data, 4 * n), "upload max_t");
	}
	std::string identification() override { return "uploaded host rays"; }
	void dont_forget_to_initialize_max_t() override {}
};

template <typename box_t, typename tri_t> class gpu_ray_bouncer : virtual public rta::bouncer {
public:
	uint w, h;
	triangle_intersection<tri_t> *gpu_last_intersection;
	gpu_ray_bouncer(uint w, uint h) : w(w), h(h), gpu_last_intersection(0) {
		void *p = 0;
		check(rta_hip_malloc(&p, (size_t)w * h * sizeof(triangle_intersection<tri_t>)),
		      "intersection buffer");
		gpu_last_intersection = (triangle_intersection<tri_t> *)p;
	}
	~gpu_ray_bouncer() override { rta_hip_free(gpu_last_intersection); }
};

template <typename box_t, typename tri_t>
class primary_intersection_collector : public gpu_ray_bouncer<box_t, tri_t> {
public:
	primary_intersection_collector(uint w, uint h) : gpu_ray_bouncer<box_t, tri_t>(w, h) {}
	bool trace_further_bounces() override { return false; }
	void bounce() override {}
	std::string identification() override { return "HIP primary intersection collector"; }
};

//! Downloads the intersections after each trace for host consumers.
template <typename box_t, typename tri_t>
class primary_intersection_downloader : public primary_intersection_collector<box_t, tri_t> {
public:
	std::vector<triangle_intersection<tri_t>> last_intersection;
	primary_intersection_downloader(uint w, uint h)
	    : primary_intersection_collector<box_t, tri_t>(w, h), last_intersection((size_t)w * h) {}
	void bounce() override {
		check(rta_hip_memcpy_dtoh(last_intersection.data(), this->gpu_last_intersection,
		                          sizeof(triangle_intersection<tri_t>) * last_intersection.size()),
		      "download intersections");
	}
	std::string identification() override { return "HIP primary intersections, host copy"; }
};

template <typename box_t, typename tri_t> class gpu_raytracer : public basic_raytracer<box_t, tri_t> {
protected:
	gpu_ray_bouncer<box_t, tri_t> *gpu_bouncer;
	gpu_ray_generator *gpu_raygen;

public:
	gpu_raytracer(rta::ray_generator *rg, rta::bouncer *b, basic_acceleration_structure<box_t, tri_t> *as)
	    : basic_raytracer<box_t, tri_t>(rg, b, as), gpu_bouncer(0), gpu_raygen(0) {
		ray_bouncer(b);
		ray_generator(rg);
	}
	void ray_bouncer(rta::bouncer *b) override {
		basic_raytracer<box_t, tri_t>::ray_bouncer(b);
		gpu_bouncer = dynamic_cast<gpu_ray_bouncer<box_t, tri_t> *>(b);
	}
	void ray_generator(rta::ray_generator *rg) override {
		basic_raytracer<box_t, tri_t>::ray_generator(rg);
		gpu_raygen = dynamic_cast<gpu_ray_generator *>(rg);
	}
	void prepare_trace() override {
		check(rta_reset_intersections((rta_hit *)gpu_bouncer->gpu_last_intersection, gpu_bouncer->w,
		                              gpu_bouncer->h, 0),
		      "rta_reset_intersections");
		check(rta_hip_device_synchronize(), "reset");
	}
};

}  // namespace hip
}  // namespace rta
