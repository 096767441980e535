// Seeded synthetic scenes and the host pinhole ray generator.
// Geometry is computed in double and rounded once to float.
#include "../../../include/rta_host.h"

#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
	z += 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

// uniform double in [0,1) with 53 random bits
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct D3 {
	double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double len(D3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline D3 normalized(D3 a) {
	double l = len(a);
	return {a.x / l, a.y / l, a.z / l};
}

inline rta_float3 f3(D3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }

void fill_tri(rta_triangle &t, D3 a, D3 b, D3 c, D3 na, D3 nb, D3 nc) {
	std::memset(&t, 0, sizeof t);
	t.a = f3(a); t.b = f3(b); t.c = f3(c);
	t.na = f3(na); t.nb = f3(nb); t.nc = f3(nc);
	t.material_index = -1;
}

// uniform direction on the unit sphere from two uniforms (inverse CDF)
D3 sphere_dir(uint64_t h1, uint64_t h2) {
	double z = 1.0 - 2.0 * unit(h1);
	double r = std::sqrt(std::max(0.0, 1.0 - z * z));
	double phi = 2.0 * M_PI * unit(h2);
	return {r * std::cos(phi), r * std::sin(phi), z};
}

// ---------------------------------------------------------- value noise

struct ValueNoise {
	uint64_t seed;
	double lattice(int64_t ix, int64_t iy) const {
		uint64_t h = mix64(seed ^ mix64((uint64_t)ix * 0x8cb92ba72f3d8dd7ull ^
		                                (uint64_t)iy * 0xd6e8feb86659fd93ull));
		return 2.0 * unit(h) - 1.0;
	}
	double operator()(double x, double y) const {
		double fx = std::floor(x), fy = std::floor(y);
		int64_t ix = (int64_t)fx, iy = (int64_t)fy;
		double tx = x - fx, ty = y - fy;
		double sx = tx * tx * (3.0 - 2.0 * tx), sy = ty * ty * (3.0 - 2.0 * ty);
		double v00 = lattice(ix, iy), v10 = lattice(ix + 1, iy);
		double v01 = lattice(ix, iy + 1), v11 = lattice(ix + 1, iy + 1);
		double a = v00 + sx * (v10 - v00);
		double b = v01 + sx * (v11 - v01);
		return a + sy * (b - a);
	}
};

}  // namespace

extern "C" {

int rta_scene_icosphere(int subdiv, float radius, const float *center, float jitter,
                        uint64_t seed, rta_triangle *out) {
	if (subdiv < 0 || subdiv > 10) return -1;
	int count = 20;
	for (int s = 0; s < subdiv; ++s) count *= 4;
	if (!out) return count;

	const double t = (1.0 + std::sqrt(5.0)) / 2.0;
	std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
	                     {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
	                     {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
	for (D3 &p : v) p = normalized(p);
	std::vector<int> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
	                      1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
	                      3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
	                      4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
	for (int s = 0; s < subdiv; ++s) {
		std::map<std::pair<int, int>, int> mid;
		auto midpoint = [&](int a, int b) {
			std::pair<int, int> key(std::min(a, b), std::max(a, b));
			auto it = mid.find(key);
			if (it != mid.end()) return it->second;
			D3 m = normalized(0.5 * (v[a] + v[b]));
			v.push_back(m);
			int id = (int)v.size() - 1;
			mid[key] = id;
			return id;
		};
		std::vector<int> nf;
		nf.reserve(f.size() * 4);
		for (size_t i = 0; i < f.size(); i += 3) {
			int a = f[i], b = f[i + 1], c = f[i + 2];
			int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
			int tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
			nf.insert(nf.end(), tri, tri + 12);
		}
		f.swap(nf);
	}
	D3 ctr = {0, 0, 0};
	if (center) ctr = {center[0], center[1], center[2]};
	std::vector<D3> pos(v.size());
	for (size_t i = 0; i < v.size(); ++i) {
		double r = radius;
		if (jitter > 0) r *= 1.0 + jitter * (2.0 * unit(mix64(seed ^ mix64(i + 1))) - 1.0);
		pos[i] = ctr + r * v[i];
	}
	for (int i = 0; i < count; ++i) {
		int a = f[3 * i], b = f[3 * i + 1], c = f[3 * i + 2];
		fill_tri(out[i], pos[a], pos[b], pos[c], v[a], v[b], v[c]);
	}
	return count;
}

int rta_scene_heightfield(int nx, int ny, float x0, float y0, float x1, float y1, float z0,
                          uint64_t seed, int octaves, float amplitude, float lacunarity,
                          float gain, float base_freq, rta_triangle *out) {
	if (nx <= 0 || ny <= 0 || octaves <= 0) return -1;
	const long long count = 2LL * nx * ny;
	if (count > 0x7fffffffLL) return -1;
	if (!out) return (int)count;
	ValueNoise vn{mix64(seed)};
	double wsum = 0, w = 1;
	for (int o = 0; o < octaves; ++o) { wsum += w; w *= gain; }
	auto height = [&](double x, double y) {
		double s = 0, amp = 1, fr = base_freq;
		for (int o = 0; o < octaves; ++o) {
			ValueNoise vo{mix64(vn.seed + (uint64_t)o)};
			s += amp * vo(fr * x, fr * y);
			amp *= gain;
			fr *= lacunarity;
		}
		return (double)z0 + (double)amplitude * s / wsum;
	};
	const int vx = nx + 1, vy = ny + 1;
	std::vector<D3> p((size_t)vx * vy), nrm((size_t)vx * vy);
	const double dx = ((double)x1 - x0) / nx, dy = ((double)y1 - y0) / ny;
	for (int j = 0; j < vy; ++j)
		for (int i = 0; i < vx; ++i) {
			double x = x0 + i * dx, y = y0 + j * dy;
			p[(size_t)j * vx + i] = {x, y, height(x, y)};
		}
	for (int j = 0; j < vy; ++j)
		for (int i = 0; i < vx; ++i) {
			int il = i > 0 ? i - 1 : i, ir = i < vx - 1 ? i + 1 : i;
			int jl = j > 0 ? j - 1 : j, jr = j < vy - 1 ? j + 1 : j;
			double hx = (p[(size_t)j * vx + ir].z - p[(size_t)j * vx + il].z) / ((ir - il) * dx);
			double hy = (p[(size_t)jr * vx + i].z - p[(size_t)jl * vx + i].z) / ((jr - jl) * dy);
			nrm[(size_t)j * vx + i] = normalized({-hx, -hy, 1.0});
		}
	long long k = 0;
	for (int j = 0; j < ny; ++j)
		for (int i = 0; i < nx; ++i) {
			size_t v00 = (size_t)j * vx + i, v10 = v00 + 1, v01 = v00 + vx, v11 = v01 + 1;
			fill_tri(out[k++], p[v00], p[v10], p[v11], nrm[v00], nrm[v10], nrm[v11]);
			fill_tri(out[k++], p[v00], p[v11], p[v01], nrm[v00], nrm[v11], nrm[v01]);
		}
	return (int)count;
}

int rta_scene_instances(int n_inst, int g, uint64_t seed, int *n_verts, float *verts,
                        uint32_t *idx) {
	if (n_inst <= 0 || g < 2 || (g & (g - 1)) != 0 || !n_verts) return -1;
	const long long tris_per = 2LL * g * g;
	if (tris_per * n_inst > 0x7fffffffLL) return -1;
	int s = 0;
	while ((8LL << (2 * s)) < tris_per) ++s;  // 8 * 4^s == 2 g^2
	const long long hf_verts = (long long)(g + 1) * (g + 1);
	const long long sp_verts = (long long)g * g + 2;
	const int side = (int)std::ceil(std::sqrt((double)n_inst));
	auto is_sphere = [&](int k) { return (mix64(seed ^ mix64(0x1157ull + (uint64_t)k)) & 1) != 0; };
	long long nv = 0;
	for (int k = 0; k < n_inst; ++k) nv += is_sphere(k) ? sp_verts : hf_verts;
	if (nv > 0x7fffffffLL) return -1;
	*n_verts = (int)nv;
	const int count = (int)(tris_per * n_inst);
	if (!verts || !idx) return count;

	// unit geodesic sphere: octahedron subdivided s times, shared vertices
	std::vector<D3> sv = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
	std::vector<uint32_t> sf = {4, 0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0,
	                            5, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3};
	for (int it = 0; it < s; ++it) {
		std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
		auto midpoint = [&](uint32_t a, uint32_t b) {
			std::pair<uint32_t, uint32_t> key(std::min(a, b), std::max(a, b));
			auto f = mid.find(key);
			if (f != mid.end()) return f->second;
			sv.push_back(normalized(0.5 * (sv[a] + sv[b])));
			const uint32_t id = (uint32_t)sv.size() - 1;
			mid[key] = id;
			return id;
		};
		std::vector<uint32_t> nf;
		nf.reserve(sf.size() * 4);
		for (size_t i = 0; i < sf.size(); i += 3) {
			const uint32_t a = sf[i], b = sf[i + 1], c = sf[i + 2];
			const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
			const uint32_t t4[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
			nf.insert(nf.end(), t4, t4 + 12);
		}
		sf.swap(nf);
	}

	long long vbase = 0, t = 0;
	for (int k = 0; k < n_inst; ++k) {
		const double cx = (double)(k % side), cy = (double)(k / side);
		const uint64_t hk = mix64(seed ^ mix64(0x2000ull + (uint64_t)k));
		float *vp = verts + 3 * vbase;
		if (is_sphere(k)) {
			const double r = 0.3 + 0.1 * unit(mix64(hk + 1));
			const D3 c = {cx + 0.5, cy + 0.5, 0.1 + r};
			for (size_t i = 0; i < sv.size(); ++i) {
				const D3 p = c + r * sv[i];
				vp[3 * i] = (float)p.x; vp[3 * i + 1] = (float)p.y; vp[3 * i + 2] = (float)p.z;
			}
			for (size_t i = 0; i < sf.size(); ++i) idx[3 * t + i] = (uint32_t)(vbase + sf[i]);
			vbase += (long long)sv.size();
		} else {
			ValueNoise vn{mix64(hk + 2)};
			const int vx = g + 1;
			for (int j = 0; j <= g; ++j)
				for (int i = 0; i <= g; ++i) {
					const double x = cx + (double)i / g, y = cy + (double)j / g;
					double hsum = 0, amp = 1, fr = 4.0, wsum = 0;
					for (int o = 0; o < 6; ++o) {
						ValueNoise vo{mix64(vn.seed + (uint64_t)o)};
						hsum += amp * vo(fr * x, fr * y);
						wsum += amp;
						amp *= 0.5;
						fr *= 2.0;
					}
					const size_t q = (size_t)j * vx + i;
					vp[3 * q] = (float)x; vp[3 * q + 1] = (float)y;
					vp[3 * q + 2] = (float)(0.2 * hsum / wsum);
				}
			uint32_t *ip = idx + 3 * t;
			for (int j = 0; j < g; ++j)
				for (int i = 0; i < g; ++i) {
					const uint32_t v00 = (uint32_t)(vbase + (long long)j * vx + i), v10 = v00 + 1,
					               v01 = v00 + (uint32_t)vx, v11 = v01 + 1;
					const uint32_t q[6] = {v00, v10, v11, v00, v11, v01};
					for (int e = 0; e < 6; ++e) *ip++ = q[e];
				}
			vbase += hf_verts;
		}
		t += tris_per;
	}
	return count;
}

int rta_scene_random_triangles(int count, uint64_t seed, float lo, float hi, float edge_min,
                               float edge_max, rta_triangle *out) {
	if (count < 0) return -1;
	if (!out) return count;
	const uint64_t s0 = mix64(seed ^ 0x5851f42d4c957f2dull);
	for (int i = 0; i < count; ++i) {
		uint64_t b = mix64(s0 + (uint64_t)i * 16);
		D3 c = {lo + (hi - lo) * unit(mix64(b + 1)), lo + (hi - lo) * unit(mix64(b + 2)),
		        lo + (hi - lo) * unit(mix64(b + 3))};
		D3 vtx[3];
		for (int k = 0; k < 3; ++k) {
			double l = edge_min + (edge_max - edge_min) * unit(mix64(b + 4 + 3 * k));
			D3 d = sphere_dir(mix64(b + 5 + 3 * k), mix64(b + 6 + 3 * k));
			vtx[k] = c + l * d;
		}
		D3 e1 = {vtx[1].x - vtx[0].x, vtx[1].y - vtx[0].y, vtx[1].z - vtx[0].z};
		D3 e2 = {vtx[2].x - vtx[0].x, vtx[2].y - vtx[0].y, vtx[2].z - vtx[0].z};
		D3 n = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
		double nl = len(n);
		if (nl > 0) n = {n.x / nl, n.y / nl, n.z / nl};
		fill_tri(out[i], vtx[0], vtx[1], vtx[2], n, n, n);
	}
	return count;
}

int rta_host_shirley_rays(float *dirs, float *origins, float *max_t, float fovy, float aspect,
                          int w, int h, const float *vd, const float *pos, const float *up,
                          float maxt) {
	if (!dirs || !origins || !max_t || !vd || !pos || !up || w <= 0 || h <= 0) return -1;
	// camera basis: identical float operations to the device generator
	float fh = (float)((double)fovy / 2.0);
	float height = tanf((float)(M_PI * (double)fh / 180.0f));
	float width = aspect * height;
	const float wl = sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
	const float wx = vd[0] / wl, wy = vd[1] / wl, wz = vd[2] / wl;
	float tx = up[1] * wz - up[2] * wy;
	float ty = up[2] * wx - up[0] * wz;
	float tz = up[0] * wy - up[1] * wx;
	const float tl = sqrtf(tx * tx + ty * ty + tz * tz);
	const float ux = tx / tl, uy = ty / tl, uz = tz / tl;
	const float vx = wy * uz - wz * uy, vy = wz * ux - wx * uz, vz = wx * uy - wy * ux;
	for (int y = 0; y < h; ++y)
		for (int x = 0; x < w; ++x) {
			const size_t id = (size_t)y * w + x;
			origins[3 * id + 0] = pos[0];
			origins[3 * id + 1] = pos[1];
			origins[3 * id + 2] = pos[2];
			max_t[id] = maxt;
			float u_s = ((float)x + 0.5f) / (float)w * 2.0f - 1.0f;
			float v_s = ((float)y + 0.5f) / (float)h * 2.0f - 1.0f;
			u_s = width * u_s;
			v_s = height * v_s;
			float dx = ux * u_s, dy = uy * u_s, dz = uz * u_s;
			dx = dx + vx * v_s;
			dy = dy + vy * v_s;
			dz = dz + vz * v_s;
			dx = dx + wx;
			dy = dy + wy;
			dz = dz + wz;
			const float l = sqrtf(dx * dx + dy * dy + dz * dz);
			dirs[3 * id + 0] = dx / l;
			dirs[3 * id + 1] = dy / l;
			dirs[3 * id + 2] = dz / l;
		}
	return 0;
}

}  // extern "C"
