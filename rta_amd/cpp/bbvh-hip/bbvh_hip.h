// bbvh-hip plugin objects (the roles of bbvh-cuda/lbvh.h:13-216 and
// bbvh-cuda/bbvh-cuda.h:79-235): the LBVH acceleration structure living in
// device memory, its constructor (the seven-stage GPU build), and the
// closest-hit child-is tracer.  All device work goes through rta_hip.h.
#pragma once

#include <sys/time.h>

#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../include/rta/hip_glue.h"
#include "../../../include/rta_host.h"

namespace rta {
namespace hip {

inline double wall_ms() {
	timeval tv;
	gettimeofday(&tv, 0);
	return (double)tv.tv_sec * 1000.0 + (double)tv.tv_usec / 1000.0;
}

//! Binary BVH in device memory: the triangles in the tree's order, the
//! nodes, the index buffer (tree position -> input triangle) and, on first
//! use, the packed traversal layout.  An LBVH (Karras) has 2n-1 nodes, its
//! triangle array is the one handed to the constructor (reordered in place)
//! and the constructor owns nodes and index buffer; a host-built median tree
//! has fewer nodes and owns device copies of all three (`owns_data`).
template <typename box_t, typename tri_t>
class binary_lbvh : public basic_acceleration_structure<box_t, tri_t> {
public:
	tri_t *triangles;
	uint n;
	rta_node *nodes;
	uint32_t *index_buffer;
	uint n_nodes;
	bool owns_data = false;
	bool owns_layout = true;   //!< false: the constructor owns pairs/tri4/flags
	std::string name = "binary lbvh (Karras) on gfx950";
	rta_node_pair *pairs = 0;
	float *tri4 = 0;
	uint32_t *flags = 0;
	uint32_t root_link = 0, root_index = 0;

	binary_lbvh(tri_t *tris, uint n, rta_node *nodes, uint32_t *index)
	    : triangles(tris), n(n), nodes(nodes), index_buffer(index), n_nodes(2 * n - 1) {}
	~binary_lbvh() override {
		if (owns_layout) {
			rta_hip_free(pairs);
			rta_hip_free(tri4);
			rta_hip_free(flags);
		}
		if (owns_data) {
			rta_hip_free(triangles);
			rta_hip_free(nodes);
			rta_hip_free(index_buffer);
		}
	}
	std::string identification() override { return name; }
	tri_t *triangle_ptr() override { return triangles; }
	int triangle_count() override { return (int)n; }
	uint node_count() const { return n_nodes; }
	tri_t *canonical_triangle_ptr() override {
		tri_t *t = new tri_t[n];
		check(rta_hip_memcpy_dtoh(t, triangles, sizeof(tri_t) * (size_t)n), "download triangles");
		return t;
	}
	void free_canonical_triangles(tri_t *data) override { delete[] data; }

	std::vector<rta_node> download_nodes() const {
		std::vector<rta_node> v(node_count());
		check(rta_hip_memcpy_dtoh(v.data(), nodes, sizeof(rta_node) * v.size()), "download nodes");
		return v;
	}

	//! derive the child-pair layout (once per build)
	void pack() {
		if (pairs) return;
		void *p = 0;
		check(rta_hip_malloc(&p, sizeof(rta_node_pair) * (size_t)node_count()), "pairs");
		pairs = (rta_node_pair *)p;
		check(rta_hip_malloc(&p, 48 * (size_t)n), "tri4");
		tri4 = (float *)p;
		check(rta_hip_malloc(&p, 4), "flags");
		flags = (uint32_t *)p;
		check(rta_bvh_pack(pairs, tri4, flags, nodes, (int)node_count(), (const rta_triangle *)triangles,
		                   (int)n, 0),
		      "rta_bvh_pack");
		rta_node root;
		check(rta_hip_memcpy_dtoh(&root, nodes, sizeof root), "root");
		root_link = root.type_left_elems;
		root_index = root.right_tris;
	}

	//! text dump in the reference's .bvh format (bbvh/bbvh.h:36-52)
	void dump_acceleration_structure(const std::string &filename) override {
		std::vector<rta_node> v = download_nodes();
		std::ofstream out(filename.c_str());
		out << "bvh\n" << v.size() << "\n";
		for (size_t i = 0; i < v.size(); ++i) {
			const rta_node &nd = v[i];
			const bool inner = (nd.type_left_elems & 1u) != 0;
			out << i << " " << (inner ? "I " : "L ") << nd.box.min.x << " " << nd.box.min.y << " "
			    << nd.box.min.z << " " << nd.box.max.x << " " << nd.box.max.y << " " << nd.box.max.z
			    << " ";
			if (inner)
				out << 0 << " " << (nd.type_left_elems >> 1) << " " << nd.right_tris << "\n";
			else {
				out << (nd.type_left_elems >> 1);
				for (uint j = 0; j < (nd.type_left_elems >> 1); ++j) out << " " << nd.right_tris + j;
				out << "\n";
			}
		}
	}
	//! text dump in the reference's .tri format (bbvh/bbvh.h:53-65)
	void dump_primitives(const std::string &filename) override {
		tri_t *t = canonical_triangle_ptr();
		std::ofstream out(filename.c_str());
		out << "tri\n" << n << "\n";
		for (uint i = 0; i < n; ++i)
			out << i << " " << t[i].a.x << " " << t[i].a.y << " " << t[i].a.z << " " << t[i].b.x << " "
			    << t[i].b.y << " " << t[i].b.z << " " << t[i].c.x << " " << t[i].c.y << " " << t[i].c.z
			    << " " << t[i].na.x << " " << t[i].na.y << " " << t[i].na.z << " " << t[i].nb.x << " "
			    << t[i].nb.y << " " << t[i].nb.z << " " << t[i].nc.x << " " << t[i].nc.y << " "
			    << t[i].nc.z << "\n";
		free_canonical_triangles(t);
	}
};

//! GPU LBVH constructor: triangles must be in device memory
//! (expects_host_triangles() == false); they become the structure's
//! triangle array.  Buffers are reused while n does not change
//! (bbvh-cuda/lbvh.h:157-162).
template <typename box_t, typename tri_t>
class lbvh_constructor : public basic_acceleration_structure_constructor<box_t, tri_t> {
	uint last_n = 0;
	void *workspace = 0;
	size_t workspace_bytes = 0;
	rta_node *d_nodes = 0;
	uint32_t *d_indices = 0;
	uint32_t *d_status = 0;
	rta_node_pair *d_pairs = 0;
	float *d_tri4 = 0;
	uint32_t *d_flags = 0;

	void free_data() {
		rta_hip_free(workspace);
		rta_hip_free(d_nodes);
		rta_hip_free(d_indices);
		rta_hip_free(d_status);
		rta_hip_free(d_pairs);
		rta_hip_free(d_tri4);
		rta_hip_free(d_flags);
		workspace = 0;
		d_nodes = 0;
		d_indices = 0;
		d_status = 0;
		d_pairs = 0;
		d_tri4 = 0;
		d_flags = 0;
	}

public:
	bool verbose = false;
	rta_lbvh_times last_times{};
	uint32_t last_status = 0;

	~lbvh_constructor() override { free_data(); }
	std::string identification() override { return "lbvh constructor (bbvh-hip)"; }
	bool expects_host_triangles() override { return false; }

	binary_lbvh<box_t, tri_t> *build(typename tri_t::input_flat_triangle_list_t *tris) override {
		const uint n = tris->triangles;
		if (n == 0) throw std::runtime_error("lbvh_constructor: cannot build 0 triangles");
		if (workspace == 0 || last_n != n) {
			free_data();
			check(rta_lbvh_workspace_bytes((int)n, &workspace_bytes), "workspace size");
			check(rta_hip_malloc(&workspace, workspace_bytes), "workspace");
			void *p = 0;
			check(rta_hip_malloc(&p, sizeof(rta_node) * (2 * (size_t)n - 1)), "nodes");
			d_nodes = (rta_node *)p;
			check(rta_hip_malloc(&p, 4 * (size_t)n), "indices");
			d_indices = (uint32_t *)p;
			check(rta_hip_malloc(&p, 4), "status");
			d_status = (uint32_t *)p;
			check(rta_hip_malloc(&p, sizeof(rta_node_pair) * (2 * (size_t)n - 1)), "pairs");
			d_pairs = (rta_node_pair *)p;
			check(rta_hip_malloc(&p, 48 * (size_t)n), "tri4");
			d_tri4 = (float *)p;
			check(rta_hip_malloc(&p, 4), "flags");
			d_flags = (uint32_t *)p;
		}
		last_n = n;
		// the build writes the packed traversal layout in its own passes
		check(rta_lbvh_build_packed((rta_triangle *)tris->triangle, (int)n, d_nodes, d_indices,
		                            d_pairs, d_tri4, d_flags, workspace, workspace_bytes,
		                            d_status, &last_times, 0),
		      "rta_lbvh_build_packed");
		check(rta_hip_memcpy_dtoh(&last_status, d_status, 4), "status");
		if (last_status & RTA_LBVH_FRONTIER_OVERFLOW)
			throw std::runtime_error("rta_lbvh_build: bottom-up frontier overflow, upper boxes not merged");
		if (verbose) {
			std::cout << "LBVH build of " << n << " triangles [ms]: init " << last_times.init_triangles
			          << ", bounds " << last_times.voxel_bounds << ", morton " << last_times.morton
			          << ", sort " << last_times.sort << ", tree " << last_times.tree
			          << ", permute " << last_times.permute << ", boxes " << last_times.boxes
			          << ", total (with the traversal layout) " << last_times.total << std::endl;
			if (last_status & RTA_LBVH_CODE_ZERO)
				std::cout << "note: a Morton code is 0 (the reference's bottom-up pass is racy there)"
				          << std::endl;
		}
		auto *bvh = new binary_lbvh<box_t, tri_t>(tris->triangle, n, d_nodes, d_indices);
		bvh->pairs = d_pairs;
		bvh->tri4 = d_tri4;
		bvh->flags = d_flags;
		bvh->owns_layout = false;
		rta_node root;
		check(rta_hip_memcpy_dtoh(&root, d_nodes, sizeof root), "root");
		bvh->root_link = root.type_left_elems;
		bvh->root_index = root.right_tris;
		return bvh;
	}
};

//! Spatial-median constructor (bbvh_constructor_using_median with
//! spatial_median, bbvh/bbvh.h:110-279), the `-b median` builder of the
//! bbvh-cuda plugin (bbvh-cuda/main.cpp:158-163, 200-235): the tree is
//! built on the host from host triangles (expects_host_triangles() is
//! true) and uploaded with the triangles in leaf order; every build yields
//! a structure that owns its device arrays.
template <typename box_t, typename tri_t>
class median_constructor : public basic_acceleration_structure_constructor<box_t, tri_t> {
public:
	bool verbose = false;
	double last_build_ms = 0;

	std::string identification() override { return "bbvh_constructor_using_median: SM (bbvh-hip)"; }
	bool expects_host_triangles() override { return true; }

	binary_lbvh<box_t, tri_t> *build(typename tri_t::input_flat_triangle_list_t *tris) override {
		const int n = (int)tris->triangles;
		if (n <= 0) throw std::runtime_error("median_constructor: cannot build 0 triangles");
		const double t0 = wall_ms();
		std::vector<rta_node> nodes((size_t)rta_host_median_node_bound(n));
		std::vector<rta_triangle> out((size_t)n);
		std::vector<uint32_t> order((size_t)n);
		const int k = rta_host_median_build((const rta_triangle *)tris->triangle, n, nodes.data(),
		                                    out.data(), order.data());
		if (k <= 0) throw std::runtime_error("rta_host_median_build failed");
		last_build_ms = wall_ms() - t0;
		if (verbose)
			std::cout << "median build of " << n << " triangles on the host: " << k << " nodes, "
			          << last_build_ms << " ms" << std::endl;
		void *pt = 0, *pn = 0, *pi = 0;
		check(rta_hip_malloc(&pt, sizeof(rta_triangle) * (size_t)n), "triangles");
		check(rta_hip_malloc(&pn, sizeof(rta_node) * (size_t)k), "nodes");
		check(rta_hip_malloc(&pi, 4 * (size_t)n), "indices");
		check(rta_hip_memcpy_htod(pt, out.data(), sizeof(rta_triangle) * (size_t)n), "upload triangles");
		check(rta_hip_memcpy_htod(pn, nodes.data(), sizeof(rta_node) * (size_t)k), "upload nodes");
		check(rta_hip_memcpy_htod(pi, order.data(), 4 * (size_t)n), "upload indices");
		auto *bvh = new binary_lbvh<box_t, tri_t>((tri_t *)pt, (uint)n, (rta_node *)pn, (uint32_t *)pi);
		bvh->n_nodes = (uint)k;
		bvh->owns_data = true;
		bvh->name = "binary bvh (spatial median, host-built) on gfx950";
		return bvh;
	}
};

enum class trace_layout { plain, packed, persistent };

//! Closest-hit child-is tracer (bbvh_gpu_cis_tracer, bbvh-cuda.h:211-235).
//! trace_rays() returns the wall time of launch + kernel + synchronisation
//! like the reference's wall_time_timer around trace_cis.
template <typename box_t, typename tri_t> class bbvh_gpu_cis_tracer : public gpu_raytracer<box_t, tri_t> {
	uint32_t *d_err = 0, *d_queue = 0;

public:
	binary_lbvh<box_t, tri_t> *bvh;
	trace_layout layout;
	rta_trace_config config;
	//! the shadow form (bbvh_gpu_cis_shadow_tracer, bbvh-cuda.h:265-291):
	//! stop at the first triangle a leaf accepts; packed/persistent only
	bool shadow;

	bbvh_gpu_cis_tracer(rta::ray_generator *gen, binary_lbvh<box_t, tri_t> *bvh, rta::bouncer *b,
	                    trace_layout layout = trace_layout::persistent,
	                    rta_trace_config cfg = rta_trace_config{64, 16, 32, 0, 0},
	                    bool shadow = false)
	    : gpu_raytracer<box_t, tri_t>(gen, b, bvh), bvh(bvh), layout(layout), config(cfg),
	      shadow(shadow) {
		if (shadow && layout == trace_layout::plain)
			throw std::runtime_error("the shadow tracer needs the packed or persistent layout");
		void *p = 0;
		check(rta_hip_malloc(&p, 8), "error words");
		d_err = (uint32_t *)p;
		d_queue = d_err + 1;
		check(rta_hip_fill_u32(d_err, 0, 1, 0), "clear error word");
	}
	bbvh_gpu_cis_tracer(const bbvh_gpu_cis_tracer &o)
	    : bbvh_gpu_cis_tracer(o.raygen, o.bvh, o.bouncer, o.layout, o.config, o.shadow) {}
	~bbvh_gpu_cis_tracer() { rta_hip_free(d_err); }

	std::string identification() override {
		return shadow ? "bbvh-hip child-is shadow tracer" : "bbvh-hip closest-hit child-is tracer";
	}
	bool supports_max_t() override { return true; }
	bbvh_gpu_cis_tracer *copy() override { return new bbvh_gpu_cis_tracer(*this); }
	void acceleration_structure(basic_acceleration_structure<box_t, tri_t> *as) override {
		bvh = dynamic_cast<binary_lbvh<box_t, tri_t> *>(as);
		gpu_raytracer<box_t, tri_t>::acceleration_structure(as);
	}

	float trace_rays() override {
		gpu_ray_generator *rg = this->gpu_raygen;
		rta_hit *hits = (rta_hit *)this->gpu_bouncer->gpu_last_intersection;
		const int n = (int)(rg->w * rg->h);
		if (layout != trace_layout::plain) bvh->pack();
		const double t0 = wall_ms();
		switch (layout) {
		case trace_layout::plain:
			check(rta_trace_cis((const rta_triangle *)bvh->triangles, (int)bvh->n, bvh->nodes,
			                    rg->gpu_origin, rg->gpu_direction, rg->gpu_maxt, (int)rg->w,
			                    (int)rg->h, hits, 0, d_err, 0),
			      "rta_trace_cis");
			break;
		case trace_layout::packed:
			check((shadow ? rta_trace_shadow_cis_packed : rta_trace_cis_packed)(
			          bvh->tri4, bvh->pairs, bvh->flags, bvh->root_link, bvh->root_index,
			          rg->gpu_origin, rg->gpu_direction, rg->gpu_maxt, n, 0, hits, d_err, 0),
			      "packed trace");
			break;
		case trace_layout::persistent:
			check((shadow ? rta_trace_shadow_cis_persistent : rta_trace_cis_persistent)(
			          bvh->tri4, bvh->pairs, bvh->flags, bvh->root_link, bvh->root_index,
			          rg->gpu_origin, rg->gpu_direction, rg->gpu_maxt, n, 0, hits, d_err, d_queue,
			          &config, 0),
			      "persistent trace");
			break;
		}
		check(rta_hip_device_synchronize(), "trace");
		const float ms = (float)(wall_ms() - t0);
		uint32_t err = 0;
		check(rta_hip_memcpy_dtoh(&err, d_err, 4), "error word");
		if (err) throw std::runtime_error("traversal stack overflow / malformed tree");
		return ms;
	}
};

}  // namespace hip
}  // namespace rta
