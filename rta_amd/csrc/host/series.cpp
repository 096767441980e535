// Host side of the reference's measurement protocol (rta/main.cpp:200-462):
// scene bounds, the spherical and axial camera series, the distinct vertices
// of a sphere .ply file, and a Wavefront .obj triangle loader for -m.
//
// The reference's vector helpers come from libmcm (included by
// librta/basic_types.h:7), which is not part of the reference tree; the
// helpers below restate them in float arithmetic (normalise = divide each
// component by sqrtf of the squared length, the textbook cross product and
// the axis-angle rotation matrix).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/rta_host.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
	g_error = msg;
	return code;
}

struct V3 {
	float x, y, z;
};

V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float length(V3 a) { return sqrtf(dot(a, a)); }
V3 normalize(V3 a) {
	const float l = length(a);
	return {a.x / l, a.y / l, a.z / l};
}
V3 cross(V3 a, V3 b) {
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
V3 load(const float *p) { return {p[0], p[1], p[2]}; }
void store(float *p, V3 v) {
	p[0] = v.x;
	p[1] = v.y;
	p[2] = v.z;
}

void merge(V3 &lo, V3 &hi, const rta_float3 &v) {
	if (v.x < lo.x) lo.x = v.x;
	if (v.x > hi.x) hi.x = v.x;
	if (v.y < lo.y) lo.y = v.y;
	if (v.y > hi.y) hi.y = v.y;
	if (v.z < lo.z) lo.z = v.z;
	if (v.z > hi.z) hi.z = v.z;
}

// ---------------------------------------------------------------- .ply

struct PlyProperty {
	std::string name;
	bool list = false;
};
struct PlyElement {
	std::string name;
	long count = 0;
	std::vector<PlyProperty> props;
};

}  // namespace

extern "C" {

const char *rta_host_last_error(void) { return g_error.c_str(); }

int rta_host_scene_bounds(const rta_triangle *tris, int n, float *bb_min, float *bb_max) {
	if (!tris || n < 1 || !bb_min || !bb_max)
		return fail(-1, "rta_host_scene_bounds: need at least one triangle and two outputs");
	V3 lo = {tris[0].a.x, tris[0].a.y, tris[0].a.z}, hi = lo;
	for (int i = 0; i < n; ++i) {
		merge(lo, hi, tris[i].a);
		merge(lo, hi, tris[i].b);
		merge(lo, hi, tris[i].c);
	}
	store(bb_min, lo);
	store(bb_max, hi);
	return 0;
}

int rta_host_series_camera(const float *vertex, const float *bb_min, const float *bb_max,
                           float distance_factor, float *pos, float *dir, float *up) {
	if (!vertex || !bb_min || !bb_max || !pos || !dir || !up)
		return fail(-1, "rta_host_series_camera: null argument");
	const V3 lo = load(bb_min), hi = load(bb_max);
	V3 diff = sub(hi, lo);
	const float diam = length(diff);
	diff = scale(diff, 0.5f);
	const V3 center = add(lo, diff);
	const V3 v = load(vertex);
	const V3 from_origin = scale(normalize(v), diam * distance_factor);
	store(pos, add(center, from_origin));
	// looks at the coordinate origin, not at the scene centre (main.cpp:364)
	const V3 d = normalize(sub(V3{0, 0, 0}, v));
	V3 u = {0, 0, 1};
	if (fabsf(dot(d, u)) > 0.8f) u = {0, 1, 0};
	const V3 right = cross(d, u);
	store(dir, d);
	store(up, cross(d, right));
	return 0;
}

int rta_host_axial_positions(const float *axis, const float *anchor, int samples, float *out) {
	if (!axis || !anchor || !out || samples < 1)
		return fail(-1, "rta_host_axial_positions: need samples >= 1 and non-null arrays");
	const V3 k = normalize(load(axis));
	// the rotated point is (anchor.x, anchor.x, anchor.x), as main.cpp:302 builds it
	const V3 a = {anchor[0], anchor[0], anchor[0]};
	for (int i = 0; i < samples; ++i) {
		const float angle = (float)((2 * 3.1416 / (double)(float)samples) * i);
		const float c = cosf(angle), s = sinf(angle), t = 1.0f - c;
		const float m[3][3] = {
		    {t * k.x * k.x + c, t * k.x * k.y - s * k.z, t * k.x * k.z + s * k.y},
		    {t * k.x * k.y + s * k.z, t * k.y * k.y + c, t * k.y * k.z - s * k.x},
		    {t * k.x * k.z - s * k.y, t * k.y * k.z + s * k.x, t * k.z * k.z + c}};
		for (int r = 0; r < 3; ++r) out[3 * i + r] = m[r][0] * a.x + m[r][1] * a.y + m[r][2] * a.z;
	}
	return 0;
}

int rta_host_ply_vertices(const char *path, float *out, int max_out) {
	if (!path) return fail(-1, "rta_host_ply_vertices: null path");
	std::ifstream in(path);
	if (!in) return fail(-1, std::string("cannot open ") + path);
	std::string line;
	if (!std::getline(in, line) || line.compare(0, 3, "ply") != 0)
		return fail(-2, std::string(path) + ": not a ply file");
	std::vector<PlyElement> elems;
	bool ascii = false;
	while (std::getline(in, line)) {
		std::istringstream ls(line);
		std::string key;
		ls >> key;
		if (key == "format") {
			std::string f;
			ls >> f;
			ascii = f == "ascii";
		} else if (key == "element") {
			PlyElement e;
			ls >> e.name >> e.count;
			elems.push_back(e);
		} else if (key == "property") {
			if (elems.empty()) return fail(-2, std::string(path) + ": property before element");
			PlyProperty p;
			std::string type;
			ls >> type;
			if (type == "list") {
				std::string ct, it;
				ls >> ct >> it;
				p.list = true;
			}
			ls >> p.name;
			elems.back().props.push_back(p);
		} else if (key == "end_header") {
			break;
		}
	}
	if (!ascii) return fail(-2, std::string(path) + ": only ascii ply files are read");
	std::vector<V3> distinct;
	for (const PlyElement &e : elems) {
		int ix = -1, iy = -1, iz = -1;
		for (size_t j = 0; j < e.props.size(); ++j) {
			if (e.props[j].name == "x") ix = (int)j;
			if (e.props[j].name == "y") iy = (int)j;
			if (e.props[j].name == "z") iz = (int)j;
		}
		const bool is_vertex = e.name == "vertex";
		if (is_vertex && (ix < 0 || iy < 0 || iz < 0))
			return fail(-2, std::string(path) + ": vertex element without x, y, z");
		for (long r = 0; r < e.count; ++r) {
			V3 v = {0, 0, 0};
			for (size_t j = 0; j < e.props.size(); ++j) {
				if (e.props[j].list) {
					double cnt;
					if (!(in >> cnt)) return fail(-2, std::string(path) + ": truncated data");
					for (long q = 0; q < (long)cnt; ++q) {
						double skip;
						if (!(in >> skip)) return fail(-2, std::string(path) + ": truncated data");
					}
					continue;
				}
				float val;
				if (!(in >> val)) return fail(-2, std::string(path) + ": truncated data");
				if ((int)j == ix) v.x = val;
				if ((int)j == iy) v.y = val;
				if ((int)j == iz) v.z = val;
			}
			if (!is_vertex) continue;
			// first occurrence of each position, in file order (main.cpp:255-267)
			bool seen = false;
			for (const V3 &d : distinct)
				if (d.x == v.x && d.y == v.y && d.z == v.z) {
					seen = true;
					break;
				}
			if (!seen) distinct.push_back(v);
		}
		if (is_vertex) break;
	}
	if (out) {
		const int k = (int)distinct.size() < max_out ? (int)distinct.size() : max_out;
		for (int i = 0; i < k; ++i) store(out + 3 * i, distinct[i]);
	}
	return (int)distinct.size();
}

int rta_host_load_obj(const char *path, rta_triangle *out, int max_out) {
	if (!path) return fail(-1, "rta_host_load_obj: null path");
	std::ifstream in(path);
	if (!in) return fail(-1, std::string("cannot open ") + path);
	std::vector<V3> pos, nrm;
	std::vector<float> tc;
	int count = 0;
	std::string line;
	long lineno = 0;
	auto index = [](long i, size_t n) -> long { return i > 0 ? i - 1 : (long)n + i; };
	while (std::getline(in, line)) {
		++lineno;
		std::istringstream ls(line);
		std::string key;
		if (!(ls >> key) || key[0] == '#') continue;
		if (key == "v" || key == "vn") {
			V3 v;
			if (!(ls >> v.x >> v.y >> v.z))
				return fail(-2, std::string(path) + ":" + std::to_string(lineno) + ": bad vector");
			(key == "v" ? pos : nrm).push_back(v);
		} else if (key == "vt") {
			float u = 0, w = 0;
			ls >> u >> w;
			tc.push_back(u);
			tc.push_back(w);
		} else if (key == "f") {
			struct Corner {
				long v, t, n;
			};
			std::vector<Corner> poly;
			std::string tok;
			while (ls >> tok) {
				Corner c = {0, 0, 0};
				long *slot[3] = {&c.v, &c.t, &c.n};
				size_t start = 0;
				for (int s = 0; s < 3 && start <= tok.size(); ++s) {
					const size_t slash = tok.find('/', start);
					const std::string part =
					    tok.substr(start, slash == std::string::npos ? std::string::npos : slash - start);
					if (!part.empty()) *slot[s] = atol(part.c_str());
					if (slash == std::string::npos) break;
					start = slash + 1;
				}
				if (c.v == 0)
					return fail(-2, std::string(path) + ":" + std::to_string(lineno) + ": bad face");
				poly.push_back(c);
			}
			// polygons become triangle fans around their first corner
			for (size_t k = 2; k < poly.size(); ++k) {
				const Corner cs[3] = {poly[0], poly[k - 1], poly[k]};
				if (out && count < max_out) {
					rta_triangle &t = out[count];
					memset(&t, 0, sizeof t);
					rta_float3 *P[3] = {&t.a, &t.b, &t.c};
					rta_float3 *N[3] = {&t.na, &t.nb, &t.nc};
					rta_float2 *T[3] = {&t.ta, &t.tb, &t.tc};
					for (int q = 0; q < 3; ++q) {
						const long vi = index(cs[q].v, pos.size());
						if (vi < 0 || vi >= (long)pos.size())
							return fail(-2, std::string(path) + ":" + std::to_string(lineno) +
							                    ": vertex index out of range");
						*P[q] = {pos[vi].x, pos[vi].y, pos[vi].z};
						if (cs[q].n) {
							const long ni = index(cs[q].n, nrm.size());
							if (ni >= 0 && ni < (long)nrm.size()) *N[q] = {nrm[ni].x, nrm[ni].y, nrm[ni].z};
						}
						if (cs[q].t) {
							const long ti = index(cs[q].t, tc.size() / 2);
							if (ti >= 0 && ti < (long)(tc.size() / 2)) {
								T[q]->x = tc[2 * ti];
								T[q]->y = tc[2 * ti + 1];
							}
						}
					}
					t.material_index = -1;
				}
				++count;
			}
		}
	}
	return count;
}

}  // extern "C"
