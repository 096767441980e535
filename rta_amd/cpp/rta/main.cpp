// rta-hip: host driver for ray tracing plugins (the role of rta/main.cpp in
// the reference): loads a plugin through the four-symbol contract, hands it
// a scene, and measures it, reporting the reference's metric
// `average rps per frame: <K> K` (rps = w*h*1000/ms of the first
// trace_rays() timing of each frame, rta/main.cpp:404-407,461).
//
// Two ways to place the camera:
//  - the reference's measurement series (rta/main.cpp:200-462, options as in
//    rta/cmdline.cpp:24-50): --sphere-file (one camera per distinct vertex of
//    a .ply sphere, at distance-factor * scene diameter from the scene
//    centre), --axis/--anchor/--samples (axial rotation series) or -p/-d/-u
//    (one positional camera); all render with fovy 45;
//  - a look-at frame loop (--pos/--target/--up/--fovy, --frames N), the
//    default when no series is selected.
//
//   rta-hip [-r WxH] [-m model.obj | --scene icosphere[:S] | terrain[:GRID:RANDOM]]
//           [--sphere-file F.ply | --axis x,y,z [--anchor x,y,z] [-s N] | -p x,y,z [-d x,y,z] [-u x,y,z]]
//           [--distance-factor F] [-f p|a|s] [-o out.ply] [-M prefix] [-R]
//           [--frames N] [--pos x,y,z] [--target x,y,z] [--up x,y,z] [--fovy F]
//           [--dump-hits FILE] [--dump-bvh FILE] [plugin] [-- plugin options]
#include <dlfcn.h>
#include <libgen.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/rta_host.h"
#include "../include/rta/hip_glue.h"
#include "../include/rta/librta.h"

using namespace rta;

namespace {

enum class Series { none, positional, axial, sphere };

struct Options {
	int w = 512, h = 512, frames = 1;
	std::string scene = "icosphere:5";
	std::string model;
	std::string plugin = "bbvh-hip";
	vec3f pos{0, 0, 3}, target{0, 0, 0}, up{0, 1, 0};
	float fovy = 60.0f;
	bool rebuild = false, camera_set = false;
	std::string dump_hits, dump_bvh;
	// measurement series (rta/cmdline.h:10-45)
	bool positional = false, axial = false, sphere = false;
	vec3f series_pos{0, 0, 0}, series_dir{0, 0, -1}, series_up{0, 1, 0};
	vec3f axis{0, 0, 1}, anchor{1, 0, 0};
	int samples = 0;
	std::string sphere_file, outfile, force, mamo_prefix;
	float distance_factor = 1.5f;
};

bool parse_vec(const char *s, vec3f &v) { return sscanf(s, "%f,%f,%f", &v.x, &v.y, &v.z) == 3; }

int usage() {
	std::cerr << "usage: rta-hip [-r WxH] [-m model.obj | --scene icosphere[:S]|terrain[:GRID:RANDOM]]\n"
	             "       [--sphere-file F.ply | --axis x,y,z [--anchor x,y,z] [-s N] | -p x,y,z "
	             "[-d x,y,z] [-u x,y,z]]\n"
	             "       [--distance-factor F] [-f p|a|s] [-o out.ply] [-M prefix] [-R]\n"
	             "       [--frames N] [--pos x,y,z] [--target x,y,z] [--up x,y,z] [--fovy F]\n"
	             "       [--dump-hits FILE] [--dump-bvh FILE] [plugin] [-- plugin options]"
	          << std::endl;
	return 2;
}

std::vector<simple_triangle> make_scene(Options &o) {
	std::vector<simple_triangle> t;
	if (!o.model.empty()) {
		const int n = rta_host_load_obj(o.model.c_str(), nullptr, 0);
		if (n < 0) throw std::runtime_error(rta_host_last_error());
		if (n == 0) throw std::runtime_error(o.model + ": no triangles");
		t.resize(n);
		rta_host_load_obj(o.model.c_str(), (rta_triangle *)t.data(), n);
		return t;
	}
	std::vector<std::string> parts;
	std::stringstream ss(o.scene);
	for (std::string p; std::getline(ss, p, ':');) parts.push_back(p);
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
		throw std::runtime_error("unknown scene " + o.scene);
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

std::string frame_name(const std::string &prefix, int i, const char *ext) {
	std::ostringstream oss;
	oss << prefix << std::setw(3) << std::setfill('0') << std::right << i << ext;
	return oss.str();
}

// The sphere .ply again, with a float property "rps" appended to every
// vertex (the timing of the camera at that position, 0 where none) and
// meta comments naming the host and the plugin's parts
// (rta/main.cpp:272-294).  ASCII files only, like the reader.
void save_ply_with_rps(const std::string &in_name, const std::string &out_name,
                       const std::vector<float> &verts, const std::vector<float> &rps,
                       const std::vector<std::string> &comments) {
	std::ifstream in(in_name.c_str());
	if (!in) throw std::runtime_error("cannot reopen " + in_name);
	std::vector<std::string> header;
	std::string line;
	while (std::getline(in, line)) {
		header.push_back(line);
		if (line.compare(0, 10, "end_header") == 0) break;
	}
	// element order and the vertex element's property count
	std::vector<std::pair<std::string, long>> elems;
	int vertex_props = 0, x = -1, y = -1, z = -1, last_vertex_prop = -1;
	std::vector<bool> list_prop;
	for (size_t i = 0; i < header.size(); ++i) {
		std::istringstream ls(header[i]);
		std::string key;
		ls >> key;
		if (key == "element") {
			std::string name;
			long n;
			ls >> name >> n;
			elems.push_back({name, n});
		} else if (key == "property" && !elems.empty() && elems.back().first == "vertex") {
			std::string type, name;
			ls >> type;
			if (type == "list") throw std::runtime_error(in_name + ": list property in vertex element");
			ls >> name;
			if (name == "x") x = vertex_props;
			if (name == "y") y = vertex_props;
			if (name == "z") z = vertex_props;
			++vertex_props;
			last_vertex_prop = (int)i;
		}
	}
	if (last_vertex_prop < 0 || x < 0 || y < 0 || z < 0)
		throw std::runtime_error(in_name + ": no vertex positions");
	std::ofstream out(out_name.c_str());
	for (size_t i = 0; i < header.size(); ++i) {
		if (header[i].compare(0, 10, "end_header") == 0)
			for (const std::string &c : comments) out << "comment " << c << "\n";
		out << header[i] << "\n";
		if ((int)i == last_vertex_prop) out << "property float rps\n";
	}
	for (const auto &e : elems) {
		for (long r = 0; r < e.second; ++r) {
			if (!std::getline(in, line)) throw std::runtime_error(in_name + ": truncated data");
			if (e.first != "vertex") {
				out << line << "\n";
				continue;
			}
			std::istringstream ls(line);
			std::vector<float> vals(vertex_props);
			for (int k = 0; k < vertex_props; ++k) ls >> vals[k];
			float value = 0.0f;
			for (size_t j = 0; j < rps.size(); ++j)
				if (verts[3 * j] == vals[x] && verts[3 * j + 1] == vals[y] && verts[3 * j + 2] == vals[z])
					value = rps[j];
			out << line << " " << value << "\n";
		}
	}
}

}  // namespace

int main(int argc, char **argv) {
	Options o;
	int i = 1;
	for (; i < argc; ++i) {
		std::string a = argv[i];
		auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
		auto vec = [&](vec3f &dst) -> bool {
			const char *v = next();
			return v && parse_vec(v, dst);
		};
		if (a == "--") { ++i; break; }
		if (a == "-r") {
			const char *v = next();
			if (!v || sscanf(v, "%dx%d", &o.w, &o.h) != 2) return usage();
		} else if (a == "--scene") {
			const char *v = next();
			if (!v) return usage();
			o.scene = v;
		} else if (a == "-m" || a == "--model") {
			const char *v = next();
			if (!v) return usage();
			o.model = v;
		} else if (a == "--frames") {
			const char *v = next();
			if (!v) return usage();
			o.frames = atoi(v);
		} else if (a == "--pos" || a == "--target" || a == "--up") {
			vec3f &dst = a == "--pos" ? o.pos : a == "--target" ? o.target : o.up;
			if (!vec(dst)) return usage();
			o.camera_set = true;
		} else if (a == "--fovy") {
			const char *v = next();
			if (!v) return usage();
			o.fovy = (float)atof(v);
			o.camera_set = true;
		} else if (a == "-p") {
			if (!vec(o.series_pos)) return usage();
			o.positional = true;
		} else if (a == "-d" || a == "--dir") {
			if (!vec(o.series_dir)) return usage();
		} else if (a == "-u") {
			if (!vec(o.series_up)) return usage();
		} else if (a == "--axis") {
			if (!vec(o.axis)) return usage();
			o.axial = true;
		} else if (a == "--anchor") {
			if (!vec(o.anchor)) return usage();
		} else if (a == "-s" || a == "--samples") {
			const char *v = next();
			if (!v) return usage();
			o.samples = atoi(v);
		} else if (a == "--sphere-file" || a == "--sphere") {
			// argp takes unique prefixes: bbvh-cuda-bench.sh writes --sphere
			const char *v = next();
			if (!v) return usage();
			o.sphere_file = v;
			o.sphere = true;
		} else if (a == "--distance-factor" || a == "--dist") {
			const char *v = next();
			if (!v) return usage();
			o.distance_factor = (float)atof(v);
		} else if (a == "-f" || a == "--force-mode") {
			const char *v = next();
			if (!v) return usage();
			o.force = v;
		} else if (a == "-o" || a == "--outfile") {
			const char *v = next();
			if (!v) return usage();
			o.outfile = v;
		} else if (a == "-M" || a == "--mamo-output") {
			const char *v = next();
			if (!v) return usage();
			o.mamo_prefix = v;
		} else if (a == "--rebuild" || a == "-R" || a == "--rebuild-bvh") {
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
	// rta/cmdline.cpp:146-155
	if (o.force == "p") o.positional = true, o.axial = false, o.sphere = false;
	if (o.force == "a") o.positional = false, o.axial = true, o.sphere = false;
	if (o.force == "s") o.positional = false, o.axial = false, o.sphere = true;
	if ((int)o.positional + (int)o.axial + (int)o.sphere > 1) {
		std::cerr << "--pos --axis and --sphere-file are exclusive." << std::endl;
		return 1;
	}
	if (o.axial && o.samples < 1) {
		std::cerr << "--axis needs --samples N (N >= 1)" << std::endl;
		return 1;
	}
	const Series series = o.sphere       ? Series::sphere
	                      : o.axial      ? Series::axial
	                      : o.positional ? Series::positional
	                                     : Series::none;
	std::vector<char *> pargv = {(char *)"plugin"};
	for (; i < argc; ++i) pargv.push_back(argv[i]);

	try {
		// sample positions first, so that a bad sphere file fails early
		std::vector<float> verts;
		if (series == Series::sphere) {
			const int n = rta_host_ply_vertices(o.sphere_file.c_str(), nullptr, 0);
			if (n < 0) throw std::runtime_error(rta_host_last_error());
			if (n == 0) throw std::runtime_error(o.sphere_file + ": no vertices");
			verts.resize(3 * (size_t)n);
			rta_host_ply_vertices(o.sphere_file.c_str(), verts.data(), n);
		} else if (series == Series::axial) {
			verts.resize(3 * (size_t)o.samples);
			if (rta_host_axial_positions(&o.axis.x, &o.anchor.x, o.samples, verts.data()) != 0)
				throw std::runtime_error(rta_host_last_error());
		}

		register_plugin_search_path("built-plugins");
		register_plugin_search_path(exe_dir() + "/../lib");
		if (!load_plugin_functions(o.plugin)) return 1;
		std::cout << "plugin: " << plugin_description() << std::endl;
		if (plugin_parse_cmdline((int)pargv.size(), pargv.data()) != 0) return 1;

		std::vector<simple_triangle> scene = make_scene(o);
		basic_flat_triangle_list<simple_triangle> ftl;
		ftl.triangles = (uint)scene.size();
		ftl.triangle = scene.data();
		std::cout << (o.model.empty() ? "scene " + o.scene : "model " + o.model) << ": "
		          << ftl.triangles << " triangles, " << o.w << "x" << o.h << " rays" << std::endl;

		rt_set set = plugin_create_rt_set(ftl, o.w, o.h);
		if (!hip::using_hip()) {
			std::cerr << "this host drives GPU plugins only" << std::endl;
			return 1;
		}
		if (!o.mamo_prefix.empty()) {
			set.as->dump_acceleration_structure(o.mamo_prefix + ".bvh");
			set.as->dump_primitives(o.mamo_prefix + ".tri");
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

		// rebuild from the original triangles (rta/main.cpp:389-400): host
		// triangles for a constructor that expects them, a device copy
		// otherwise
		auto rebuild = [&]() {
			auto *ctor = set.basic_ctor<box_t, tri_t>();
			basic_flat_triangle_list<tri_t> dl;
			dl.triangles = (uint)scene.size();
			if (ctor->expects_host_triangles()) {
				dl.triangle = scene.data();
			} else {
				void *p = 0;
				hip::check(rta_hip_malloc(&p, sizeof(tri_t) * scene.size()), "triangles");
				hip::check(rta_hip_memcpy_htod(p, scene.data(), sizeof(tri_t) * scene.size()),
				           "upload");
				dl.triangle = (tri_t *)p;
			}
			basic_acceleration_structure<box_t, tri_t> *as = ctor->build(&dl);
			delete set.as;
			set.as = as;
			rt->acceleration_structure(as);
		};

		if (series == Series::none) {
			vec3f dir{o.target.x - o.pos.x, o.target.y - o.pos.y, o.target.z - o.pos.z};
			cam->setup(&o.pos, &dir, &o.up, o.fovy);
			double sum = 0;
			for (int f = 0; f < o.frames; ++f) {
				if (o.rebuild) rebuild();
				if (!o.mamo_prefix.empty()) {
					cam->generate_rays();
					cam->dump_rays(frame_name(o.mamo_prefix, f, ".ray"));
				}
				rt->trace();
				const double ms = rt->timings.front();
				const double rps = (double)o.w * o.h * 1000.0 / ms;
				sum += rps;
				std::cout << "frame " << f << ": " << ms << " ms, " << rps / 1e6 << " Mrays/s"
				          << std::endl;
			}
			std::cout << "average rps per frame: " << sum / o.frames / 1000.0 << " K" << std::endl;
		} else {
			// directional_analysis_pass::run (rta/main.cpp:352-462)
			std::cout << (series == Series::sphere ? "running spherical series..."
			              : series == Series::axial ? "running axial series..."
			                                        : "running pos series...")
			          << std::endl;
			if (series == Series::sphere)
				std::cout << "got " << verts.size() / 3 << " distinct vertices" << std::endl;
			float bb_min[3], bb_max[3];
			if (rta_host_scene_bounds((const rta_triangle *)scene.data(), (int)scene.size(), bb_min,
			                          bb_max) != 0)
				throw std::runtime_error(rta_host_last_error());
			const int count = series == Series::positional ? 1 : (int)(verts.size() / 3);
			std::vector<float> timings(count);
			float sum = 0;
			for (int k = 0; k < count; ++k) {
				vec3f pos, dir, up;
				if (series == Series::positional) {
					pos = o.series_pos;
					dir = o.series_dir;
					up = o.series_up;
				} else if (rta_host_series_camera(&verts[3 * (size_t)k], bb_min, bb_max,
				                                  o.distance_factor, &pos.x, &dir.x, &up.x) != 0) {
					throw std::runtime_error(rta_host_last_error());
				}
				cam->setup(&pos, &dir, &up, 45);
				cam->generate_rays();
				if (!o.mamo_prefix.empty()) cam->dump_rays(frame_name(o.mamo_prefix, k, ".ray"));
				if (o.rebuild) rebuild();
				rt->trace();
				const float first = rt->timings.front();
				const float rps = o.w * o.h * (1000.0f / first);
				sum += rps;
				timings[k] = rps;
				std::cout << "." << std::flush;
			}
			std::cout << "\naverage rps per frame: " << (sum / count) / 1000.0f << " K" << std::endl;
			if (!o.outfile.empty() && series == Series::sphere) {
				char host[65] = {0};
				gethostname(host, 64);
				save_ply_with_rps(o.sphere_file, o.outfile, verts, timings,
				                  {"meta host: " + std::string(host),
				                   "meta accel-struct: " + set.as->identification(),
				                   "meta accel-struct-ctor: " + set.ctor->identification(),
				                   "meta ray-tracer: " + set.rt->identification(),
				                   "meta bouncer: " + bouncer->identification()});
			}
		}

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
