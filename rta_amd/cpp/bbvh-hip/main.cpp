// bbvh-hip plugin entry points: the four C symbols of the plugin contract
// (bbvh-cuda/main.cpp:347-513 plays this role for the CUDA plugin).
//
// Options (after `--` on the host's command line):
//   -v             verbose (per-stage build times)
//   -b median|lbvh spatial-median tree built on the host and uploaded
//                  (the default, as in bbvh-cuda/main.cpp:28; any value
//                  other than lbvh selects it, :88) or the GPU LBVH build
//   -t cis         child-is traversal (the only traversal of this plugin)
//   -l cpu|2f4     accepted for compatibility; both byte images are one
//   -s             shadow tracer: stop at the first accepted triangle
//                  (bbvh-cuda/main.cpp:46,93; packed/persistent only)
//   --trav plain|packed|persistent   traversal kernel (default persistent)
//   --block N --refill N --steps N --leaf-min N   persistent launch shape
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <iostream>
#include <string>

#include "../include/rta/librta.h"
#include "bbvh_hip.h"

using namespace rta;

namespace {

struct Cmdline {
	bool verbose = false;
	bool lbvh = false;   // -b lbvh; the median builder otherwise
	hip::trace_layout layout = hip::trace_layout::persistent;
	rta_trace_config config{64, 16, 32, 0, 0};
	bool shadow = false;
};
Cmdline cmdline;
char plugin_name[] = "bbvh-hip";

int usage(const char *msg, const char *arg) {
	std::cerr << "bbvh-hip: " << msg << (arg ? arg : "") << std::endl;
	std::cerr << "usage: bbvh-hip [-v] [-s] [-b median|lbvh] [-t cis] [-l cpu|2f4] "
	             "[--trav plain|packed|persistent] [--block N] [--refill N] [--steps N] [--leaf-min N]"
	          << std::endl;
	return 1;
}

}  // namespace

extern "C" {

char *description() { return (char *)"bbvh-hip: binary LBVH built and traced on MI355X (gfx950)"; }

void initialize() { hip::plugin_with_hip_loaded = true; }

int parse_cmdline(int argc, char **argv) {
	cmdline = Cmdline();
	if (argc > 0) argv[0] = plugin_name;
	for (int i = 1; i < argc; ++i) {
		std::string a = argv[i];
		auto value = [&](const char *what) -> const char * {
			if (i + 1 >= argc) return nullptr;
			(void)what;
			return argv[++i];
		};
		if (a == "-v" || a == "--verbose") cmdline.verbose = true;
		else if (a == "-b" || a == "--bvh") {
			const char *v = value("bvh");
			if (!v) return usage("missing value for ", a.c_str());
			cmdline.lbvh = std::string(v) == "lbvh";
		} else if (a == "-t" || a == "--trav") {
			const char *v = value("trav");
			if (!v) return usage("missing value for ", a.c_str());
			std::string s = v;
			if (s == "cis" || s == "persistent") cmdline.layout = hip::trace_layout::persistent;
			else if (s == "packed") cmdline.layout = hip::trace_layout::packed;
			else if (s == "plain") cmdline.layout = hip::trace_layout::plain;
			else return usage("unknown traversal ", v);
		} else if (a == "-l" || a == "--layout") {
			const char *v = value("layout");
			if (!v || (std::string(v) != "cpu" && std::string(v) != "2f4"))
				return usage("layout must be cpu or 2f4, got ", v);
		} else if (a == "--block" || a == "--refill" || a == "--steps" || a == "--leaf-min") {
			const char *v = value("n");
			if (!v) return usage("missing value for ", a.c_str());
			const int x = atoi(v);
			if (a == "--block") cmdline.config.block = x;
			else if (a == "--refill") cmdline.config.refill_min = x;
			else if (a == "--leaf-min") cmdline.config.leaf_min = x;
			else cmdline.config.steps = x;
		} else if (a == "-s" || a == "--shadow") {
			cmdline.shadow = true;
		} else if (a == "-A" || a == "--ailabox") {
			return usage("not provided by this plugin: ", a.c_str());
		} else {
			return usage("unknown option ", a.c_str());
		}
	}
	if (cmdline.shadow && cmdline.layout == hip::trace_layout::plain)
		return usage("the shadow tracer needs --trav packed or persistent", nullptr);
	return 0;
}

rt_set create_rt_set(basic_flat_triangle_list<simple_triangle> &triangle_lists, int w, int h) {
	typedef simple_aabb box_t;
	typedef simple_triangle tri_t;
	rt_set set;
	hip::binary_lbvh<box_t, tri_t> *bvh = 0;
	if (!cmdline.lbvh) {
		// the median tree is built from the host triangles and uploaded
		// (bbvh-cuda/main.cpp:158-163, 200-235)
		auto *ctor = new hip::median_constructor<box_t, tri_t>;
		ctor->verbose = cmdline.verbose;
		bvh = ctor->build(&triangle_lists);
		set.ctor = ctor;
	} else {
		// the LBVH build runs on device triangles (bbvh-cuda/main.cpp:466-470)
		basic_flat_triangle_list<tri_t> ftl;
		void *p = 0;
		hip::check(rta_hip_malloc(&p, sizeof(tri_t) * (size_t)triangle_lists.triangles),
		           "triangles");
		hip::check(rta_hip_memcpy_htod(p, triangle_lists.triangle,
		                               sizeof(tri_t) * (size_t)triangle_lists.triangles),
		           "upload triangles");
		ftl.triangle = (tri_t *)p;
		ftl.triangles = triangle_lists.triangles;
		auto *ctor = new hip::lbvh_constructor<box_t, tri_t>;
		ctor->verbose = cmdline.verbose;
		bvh = ctor->build(&ftl);
		set.ctor = ctor;
	}
	set.as = bvh;
	set.rt = new hip::bbvh_gpu_cis_tracer<box_t, tri_t>(0, bvh, 0, cmdline.layout, cmdline.config,
	                                                   cmdline.shadow);
	set.rgen = new hip::cam_ray_generator_shirley(w, h);
	return set;
}

}  // extern "C"
