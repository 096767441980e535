// rta-hip: host driver for ray tracing plugins (the role of rta/main.cpp in
// the reference): loads a plugin through the four-symbol contract, hands it
// a synthetic scene, and runs the frame loop, reporting the reference's
// metric `average rps per frame: <K> K` (rps = w*h*1000/ms of the first
// trace_rays() timing of each frame, rta/main.cpp:404-407,461).
//
//   rta-hip [-r WxH] [--scene icosphere[:S] | terrain[:GRID:RANDOM]]
//           [--frames N] [--pos x,y,z] [--target x,y,z] [--up x,y,z]
//           [--fovy F] [--rebuild] [--dump-hits FILE] [--dump-bvh FILE]
//           [plugin] [-- plugin options]
#include <dlfcn.h>
#include <libgen.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/rta_host.h"
#include "../include/rta/hip_glue.h"
#include "../include/rta/librta.h"

using namespace rta;

namespace {

struct Options {
	int w = 512, h = 512, frames = 1;
	std::string scene = "icosphere:5";
	std::string plugin = "bbvh-hip";
	vec3f pos{0, 0, 3}, target{0, 0, 0}, up{0, 1, 0};
	float fovy = 60.0f;
	bool rebuild = false, camera_set = false;
	std::string dump_hits, dump_bvh;
	std::vector<std::string> plugin_args;
};

bool parse_vec(const char *s, vec3f &v) { return sscanf(s, "%f,%f,%f", &v.x, &v.y, &v.z) == 3; }

int usage() {
	std::cerr << "usage: rta-hip [-r WxH] [--scene icosphere[:S]|terrain[:GRID:RANDOM]] "
	             "[--frames N] [--pos x,y,z] [--target x,y,z] [--up x,y,z] [--fovy F] "
	             "[--rebuild] [--dump-hits FILE] [--dump-bvh FILE] [plugin] [-- plugin options]"
	          << std::endl;
	return 2;
}

std::vector<simple_triangle> make_scene(const std::string &spec, Options &o) {
	std::vector<std::string> parts;
	std::stringstream ss(spec);
	for (std::string p; std::getline(ss, p, ':');) parts.push_back(p);
	std::vector<simple_triangle> t;
	if (parts[0] == "icosphere") {
		int s = parts.size() > 1 ? atoi(parts[1].c_str()) : 5;
		int n = rta_scene_icosphere(s, 1.0f, nullptr, 0.0f, 0, nullptr);
		t.resize(n);
		rta_scene_icosphere(s, 1.0f, nullptr, 0.0f, 0, (rta_triangle *)t.data());
	} else if (parts[0] == "terrain") {
		int grid = parts.size() > 1 ? atoi(parts[1].c_str()) : 1024;
		int rnd = parts.size() > 2 ? atoi(parts[2].c_str()) : 200000;
		int nh = rta_scene_heightfield(grid, grid, 0, 0, 1, 1, 0, 7, 6, 0.2f, 2.0f, 0.5f, 4.0f, nullptr);
		t.resize((size_t)nh + rnd);
		rta_scene_heightfield(grid, grid, 0, 0, 1, 1, 0, 7, 6, 0.2f, 2.0f, 0.5f, 4.0f,
		                      (rta_triangle *)t.data());
		rta_scene_random_triangles(rnd, 7 + 1000, 0, 1, 0.001f, 0.01f, (rta_triangle *)t.data() + nh);
		if (!o.camera_set) {
			o.pos = {0.5f, 0.05f, 0.6f};
			o.target = {0.5f, 0.55f, 0.0f};
			o.up = {0, 0, 1};
			o.fovy = 35.0f;
		}
	} else {
		throw std::runtime_error("unknown scene " + spec);
	}
	return t;
}

std::string exe_dir() {
	char buf[PATH_MAX];
	ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
	if (len <= 0) return ".";
	buf[len] = 0;
	return std::string(dirname(buf));
}

}  // namespace

int main(int argc, char **argv) {
	Options o;
	int i = 1;
	for (; i < argc; ++i) {
		std::string a = argv[i];
		auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
		if (a == "--") { ++i; break; }
		if (a == "-r") {
			const char *v = next();
			if (!v || sscanf(v, "%dx%d", &o.w, &o.h) != 2) return usage();
		} else if (a == "--scene") {
			const char *v = next();
			if (!v) return usage();
			o.scene = v;
		} else if (a == "--frames") {
			const char *v = next();
			if (!v) return usage();
			o.frames = atoi(v);
		} else if (a == "--pos" || a == "--target" || a == "--up") {
			const char *v = next();
			vec3f &dst = a == "--pos" ? o.pos : a == "--target" ? o.target : o.up;
			if (!v || !parse_vec(v, dst)) return usage();
			o.camera_set = true;
		} else if (a == "--fovy") {
			const char *v = next();
			if (!v) return usage();
			o.fovy = (float)atof(v);
			o.camera_set = true;
		} else if (a == "--rebuild") {
			o.rebuild = true;
		} else if (a == "--dump-hits") {
			const char *v = next();
			if (!v) return usage();
			o.dump_hits = v;
		} else if (a == "--dump-bvh") {
			const char *v = next();
			if (!v) return usage();
			o.dump_bvh = v;
		} else if (a[0] == '-') {
			return usage();
		} else {
			o.plugin = a;
		}
	}
	std::vector<char *> pargv = {(char *)"plugin"};
	for (; i < argc; ++i) pargv.push_back(argv[i]);

	try {
		register_plugin_search_path("built-plugins");
		register_plugin_search_path(exe_dir() + "/../lib");
		if (!load_plugin_functions(o.plugin)) return 1;
		std::cout << "plugin: " << plugin_description() << std::endl;
		if (plugin_parse_cmdline((int)pargv.size(), pargv.data()) != 0) return 1;

		std::vector<simple_triangle> scene = make_scene(o.scene, o);
		basic_flat_triangle_list<simple_triangle> ftl;
		ftl.triangles = (uint)scene.size();
		ftl.triangle = scene.data();
		std::cout << "scene " << o.scene << ": " << ftl.triangles << " triangles, " << o.w << "x"
		          << o.h << " rays" << std::endl;

		rt_set set = plugin_create_rt_set(ftl, o.w, o.h);
		if (!hip::using_hip()) {
			std::cerr << "this host drives GPU plugins only" << std::endl;
			return 1;
		}
		typedef simple_aabb box_t;
		typedef simple_triangle tri_t;
		auto *bouncer = new hip::primary_intersection_downloader<box_t, tri_t>(o.w, o.h);
		set.bouncer = bouncer;
		basic_raytracer<box_t, tri_t> *rt = set.basic_rt<box_t, tri_t>();
		if (!rt) throw std::runtime_error("plugin tracer is not a basic_raytracer<simple_aabb, simple_triangle>");
		rt->ray_bouncer(bouncer);
		rt->ray_generator(set.rgen);
		auto *cam = dynamic_cast<rta::cam_ray_generator_shirley *>(set.rgen);
		if (!cam) throw std::runtime_error("plugin ray generator is not a pinhole camera");
		vec3f dir{o.target.x - o.pos.x, o.target.y - o.pos.y, o.target.z - o.pos.z};
		cam->setup(&o.pos, &dir, &o.up, o.fovy);

		double sum = 0;
		for (int f = 0; f < o.frames; ++f) {
			if (o.rebuild) {
				// rebuild per frame from the original triangles (rta/main.cpp:389-400)
				auto *ctor = set.basic_ctor<box_t, tri_t>();
				void *p = 0;
				hip::check(rta_hip_malloc(&p, sizeof(tri_t) * scene.size()), "triangles");
				hip::check(rta_hip_memcpy_htod(p, scene.data(), sizeof(tri_t) * scene.size()),
				           "upload");
				basic_flat_triangle_list<tri_t> dl;
				dl.triangles = (uint)scene.size();
				dl.triangle = (tri_t *)p;
				basic_acceleration_structure<box_t, tri_t> *as = ctor->build(&dl);
				delete set.as;
				set.as = as;
				rt->acceleration_structure(as);
			}
			rt->trace();
			const double ms = rt->timings.front();
			const double rps = (double)o.w * o.h * 1000.0 / ms;
			sum += rps;
			std::cout << "frame " << f << ": " << ms << " ms, " << rps / 1e6 << " Mrays/s" << std::endl;
		}
		std::cout << "average rps per frame: " << sum / o.frames / 1000.0 << " K" << std::endl;

		if (!o.dump_hits.empty()) {
			std::ofstream out(o.dump_hits.c_str(), std::ios::binary);
			out.write((const char *)bouncer->last_intersection.data(),
			          sizeof(triangle_intersection<tri_t>) * bouncer->last_intersection.size());
		}
		if (!o.dump_bvh.empty()) set.as->dump_acceleration_structure(o.dump_bvh);
	} catch (const std::exception &e) {
		std::cerr << "rta-hip: " << e.what() << std::endl;
		return 1;
	}
	return 0;
}
