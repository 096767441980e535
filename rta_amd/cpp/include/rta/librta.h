// rt_set and the plugin contract (the roles of librta/librta.h:10-27 and
// librta/plugins.h:15-18): a plugin shared object exports exactly four C
// symbols, `description`, `initialize`, `parse_cmdline` and
// `create_rt_set`; the host loads it with dlopen and resolves them with
// dlsym.  rt_set carries the five polymorphic objects the plugin creates.
#pragma once

#include <string>

#include "raytrav.h"

namespace rta {

struct rt_set {
	acceleration_structure *as;
	acceleration_structure_constructor *ctor;
	rta::bouncer *bouncer;
	raytracer *rt;
	ray_generator *rgen;
	rt_set() : as(0), ctor(0), bouncer(0), rt(0), rgen(0) {}

	template <typename box_t, typename tri_t> basic_raytracer<box_t, tri_t> *basic_rt() {
		return dynamic_cast<basic_raytracer<box_t, tri_t> *>(rt);
	}
	template <typename box_t, typename tri_t> basic_acceleration_structure<box_t, tri_t> *basic_as() {
		return dynamic_cast<basic_acceleration_structure<box_t, tri_t> *>(as);
	}
	template <typename box_t, typename tri_t>
	basic_acceleration_structure_constructor<box_t, tri_t> *basic_ctor() {
		return dynamic_cast<basic_acceleration_structure_constructor<box_t, tri_t> *>(ctor);
	}
};

// the four plugin entry points, resolved by load_plugin_functions
extern char *(*plugin_description)();
extern int (*plugin_parse_cmdline)(int argc, char **argv);
extern void (*plugin_initialize)();
extern rt_set (*plugin_create_rt_set)(basic_flat_triangle_list<simple_triangle> &, int, int);

//! Directories searched (most recent first) for a plugin given by name.
void register_plugin_search_path(const std::string &path);
//! Load a plugin by name or path; tries <name>, <name>.so and <name>.so.so in
//! the search paths, then dlopen(<name>) itself.  Calls the plugin's
//! initialize() on success.  Returns false (and prints why) on failure.
bool load_plugin_functions(const std::string &plugin_name);

}  // namespace rta
