// librta-role host library: the plugin loader (the role of
// librta/plugins.cpp:8-97), the GPU-plugin flag, and the host pinhole ray
// generator.
#include <dlfcn.h>
#include <stdio.h>

#include <iostream>
#include <list>
#include <string>

#include "../../../include/rta_host.h"
#include "../include/rta/hip_glue.h"
#include "../include/rta/librta.h"

namespace rta {

char *(*plugin_description)() = 0;
int (*plugin_parse_cmdline)(int argc, char **argv) = 0;
void (*plugin_initialize)() = 0;
rt_set (*plugin_create_rt_set)(basic_flat_triangle_list<simple_triangle> &, int, int) = 0;

namespace {

std::list<std::string> &search_paths() {
	static std::list<std::string> paths;
	return paths;
}

template <typename T> bool resolve(const char *name, T &to, void *handle) {
	dlerror();
	to = (T)dlsym(handle, name);
	const char *error = dlerror();
	if (error != NULL || to == 0) {
		fprintf(stderr, "%s\n", error ? error : name);
		return false;
	}
	return true;
}

void *find_lib(const std::string &filename) {
	for (std::string path : search_paths()) {
		if (!path.empty() && path[path.size() - 1] != '/') path += "/";
		void *h = dlopen((path + filename).c_str(), RTLD_LAZY);
		if (h) return h;
	}
	return 0;
}

bool ends_with(const std::string &s, const std::string &suffix) {
	return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

}  // namespace

void register_plugin_search_path(const std::string &path) { search_paths().push_front(path); }

bool load_plugin_functions(const std::string &plugin_name) {
	if (plugin_name.empty()) {
		std::cerr << "plugin loader: empty plugin name" << std::endl;
		return false;
	}
	const bool absolute = plugin_name[0] == '/';
	void *handle = 0;
	if (!absolute) {
		handle = find_lib(plugin_name);
		if (!handle && !ends_with(plugin_name, ".so")) handle = find_lib(plugin_name + ".so");
		if (!handle) handle = find_lib(plugin_name + ".so.so");
	}
	if (!handle) handle = dlopen(plugin_name.c_str(), RTLD_LAZY);
	if (!handle) {
		std::cerr << "plugin loader: no library for plugin '" << plugin_name << "'; searched:"
		          << std::endl;
		for (const std::string &p : search_paths()) std::cerr << "  " << p << std::endl;
		std::cerr << "last dlopen error: " << dlerror() << std::endl;
		return false;
	}
	if (!resolve("description", plugin_description, handle)) return false;
	if (!resolve("initialize", plugin_initialize, handle)) return false;
	if (!resolve("create_rt_set", plugin_create_rt_set, handle)) return false;
	if (!resolve("parse_cmdline", plugin_parse_cmdline, handle)) return false;
	plugin_initialize();
	return true;
}

void cam_ray_generator_shirley::generate_rays() {
	// the generator writes separate origin / direction arrays; the rays are
	// stored interleaved per pixel (origin, direction)
	const size_t n = (size_t)res_x() * res_y();
	std::vector<vec3f> o(n), d(n);
	const int rc = rta_host_shirley_rays(&d[0].x, &o[0].x, ray_max_t.data, fovy, aspect, res_x(),
	                                     res_y(), &dir.x, &position.x, &up.x, FLT_MAX);
	if (rc != 0) throw std::runtime_error("rta_host_shirley_rays failed");
	for (size_t i = 0; i < n; ++i) {
		raydata.data[2 * i] = o[i];
		raydata.data[2 * i + 1] = d[i];
	}
}

namespace hip {
bool plugin_with_hip_loaded = false;
bool using_hip() { return plugin_with_hip_loaded; }
}  // namespace hip

}  // namespace rta
