// Seeded content for the benchmark configurations: a voxel-grid BSP map with portals, materials
// with value-noise textures, tangent-space normal maps and mip chains, and lightmaps.
//
// The reference builds maps offline with map_compiler (Quake MAP -> CSG -> BSP, map_compiler/main.rs)
// and lights them with lightmapper (lightmaps_builder.rs). Neither runs here, so this file writes the
// same data layout directly: one BSP leaf per grid cell, one quad per empty/solid cell face (facing
// into the empty cell), one portal per pair of adjacent empty cells, an axis-aligned k-d split tree
// as the BSP, and a direct-light lightmapper with voxel-grid shadow rays.
#include <algorithm>
#include <cstring>
#include <new>
#include <random>

#include "host_internal.h"

namespace swh
{

// ---------------------------------------------------------------------------------------------
// 4x4 matrices with cgmath's conventions
// ---------------------------------------------------------------------------------------------
M4 identity()
{
	M4 m;
	std::memset(&m, 0, sizeof(m));
	m.c[0].x = m.c[1].y = m.c[2].z = m.c[3].w = 1.0f;
	return m;
}

M4 mul(const M4& a, const M4& b)
{
	M4 r;
	for(int col = 0; col < 4; ++col)
		for(int row = 0; row < 4; ++row)
			r.at(row, col) = dot(a.row(row), b.c[col]);
	return r;
}

V4 mul(const M4& m, V4 v) { return v4(dot(m.row(0), v), dot(m.row(1), v), dot(m.row(2), v), dot(m.row(3), v)); }

M4 transpose(const M4& m)
{
	M4 r;
	for(int i = 0; i < 4; ++i)
		for(int j = 0; j < 4; ++j)
			r.at(i, j) = m.at(j, i);
	return r;
}

// Adjugate / determinant inverse.
bool invert(const M4& m, M4& out)
{
	float a[16];
	for(int c = 0; c < 4; ++c)
		for(int r = 0; r < 4; ++r)
			a[c * 4 + r] = m.at(r, c);
	float inv[16];
	inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
	inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
	inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
	inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
	inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
	inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
	inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
	inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
	inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
	inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
	inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
	inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
	inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
	inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
	inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
	inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
	const float det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
	if(det == 0.0f || !std::isfinite(det))
		return false;
	const float inv_det = 1.0f / det;
	for(int c = 0; c < 4; ++c)
		for(int r = 0; r < 4; ++r)
			out.at(r, c) = inv[c * 4 + r] * inv_det;
	return true;
}

M4 from_nonuniform_scale(float x, float y, float z)
{
	M4 m = identity();
	m.c[0].x = x;
	m.c[1].y = y;
	m.c[2].z = z;
	return m;
}

M4 from_translation(V3 t)
{
	M4 m = identity();
	m.c[3] = v4(t.x, t.y, t.z, 1.0f);
	return m;
}

uint32_t find_leaf(const swh_map& m, V3 p)
{
	uint32_t index = uint32_t(m.nodes.size() - 1);
	for(;;)
	{
		if(index >= FIRST_LEAF_INDEX)
			return index - FIRST_LEAF_INDEX;
		const Node& n = m.nodes[index];
		index = dot(n.plane.vec, p) - n.plane.dist >= 0.0f ? n.children[0] : n.children[1];
	}
}

} // namespace swh

using namespace swh;

namespace
{

const float CELL = 128.0f;

struct Rng
{
	std::mt19937_64 g;
	explicit Rng(uint64_t seed) : g(seed) {}
	float uniform(float a, float b) { return a + (b - a) * float(double(g() >> 11) * (1.0 / 9007199254740992.0)); }
	uint32_t below(uint32_t n) { return n ? uint32_t(g() % n) : 0; }
};

uint32_t hash32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

// Tileable value noise in [0, 1): lattice of `period` cells with smoothstep interpolation.
float value_noise(float x, float y, uint32_t period, uint32_t seed)
{
	const float fx = std::floor(x), fy = std::floor(y);
	const int ix = int(fx), iy = int(fy);
	const float tx = x - fx, ty = y - fy;
	auto lat = [&](int a, int b) {
		const uint32_t pa = uint32_t(((a % int(period)) + int(period)) % int(period));
		const uint32_t pb = uint32_t(((b % int(period)) + int(period)) % int(period));
		return float(hash32(pa * 73856093u ^ pb * 19349663u ^ seed * 83492791u) >> 8) / 16777216.0f;
	};
	const float sx = tx * tx * (3.0f - 2.0f * tx), sy = ty * ty * (3.0f - 2.0f * ty);
	const float a = lat(ix, iy) + (lat(ix + 1, iy) - lat(ix, iy)) * sx;
	const float b = lat(ix, iy + 1) + (lat(ix + 1, iy + 1) - lat(ix, iy + 1)) * sx;
	return a + (b - a) * sy;
}

float fractal_noise(float x, float y, uint32_t size, uint32_t seed)
{
	float sum = 0.0f, weight = 0.5f, norm = 0.0f;
	for(uint32_t cell = std::max(size / 4, 2u), o = 0; cell >= 2 && o < 5; cell /= 2, ++o)
	{
		sum += weight * value_noise(x / float(cell), y / float(cell), size / cell, seed + o * 1013u);
		norm += weight;
		weight *= 0.5f;
	}
	return sum / norm;
}

uint8_t to_u8(float v) { return uint8_t(std::min(255.0f, std::max(0.0f, std::round(v)))); }

} // namespace

// =============================================================================================
// textures.rs restatement: PackedNormalRoughness::pack, make_texture, build_texture_mip(s)
// =============================================================================================
namespace
{

// fast_math.rs:64-78: cvtps2dq (round half to even) then signed saturation to i8.
int32_t pack_f32x4_into_bytes(const float v[4], const float scale[4])
{
	int32_t res = 0;
	for(int i = 0; i < 4; ++i)
	{
		const float s = v[i] * scale[i];
		long q = std::lrint(s); // default rounding mode: to nearest, ties to even
		if(!(s >= -2147483648.0f && s < 2147483648.0f))
			q = INT32_MIN;
		q = std::max(-128L, std::min(127L, q));
		res |= int32_t(uint32_t(q) & 0xFFu) << (8 * i);
	}
	return res;
}

int32_t pack_normal_roughness(V3 n, float roughness)
{
	const float v[4] = {n.x, n.y, n.z, roughness};
	const float s[4] = {127.0f, 127.0f, 127.0f, 127.0f};
	return pack_f32x4_into_bytes(v, s);
}

void unpack_normal_roughness(int32_t p, V3& n, float& r)
{
	const float k = 1.0f / 127.0f;
	n = v3(float(int8_t(p & 0xFF)) * k, float(int8_t((p >> 8) & 0xFF)) * k, float(int8_t((p >> 16) & 0xFF)) * k);
	r = float(int8_t((p >> 24) & 0xFF)) * k;
}

const float MIN_VALID_ROUGHNESS = 1.0f / (127.0f - 1.0f);

// textures.rs:420-431
V3 renormalize_normal(V3 n)
{
	const float len = length(n);
	if(len <= 0.000001f)
		return v3(0.0f, 0.0f, 1.0f);
	return n / len;
}

// textures.rs:126-158 for one texel
int32_t make_texel_normal_roughness(uint32_t normal_rgb, bool has_normal_map, float roughness)
{
	V3 n = v3(0.0f, 0.0f, 1.0f);
	if(has_normal_map)
	{
		const int32_t zero_level = 128;
		n = renormalize_normal(v3(float(int32_t((normal_rgb >> 16) & 0xFF) - zero_level),
			float(int32_t((normal_rgb >> 8) & 0xFF) - zero_level), float(int32_t(normal_rgb & 0xFF) - zero_level)));
	}
	const float r = std::min(std::max(roughness, MIN_VALID_ROUGHNESS), 1.0f);
	return pack_normal_roughness(n, r);
}

// color.rs:51-71
uint32_t average_4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3)
{
	uint32_t s[4] = {0, 0, 0, 0};
	for(uint32_t c : {c0, c1, c2, c3})
	{
		s[0] += c & 0x000000FFu;
		s[1] += c & 0x0000FF00u;
		s[2] += c & 0x00FF0000u;
		s[3] += (c & 0xFF000000u) >> 2;
	}
	return (s[0] >> 2) | ((s[1] >> 2) & 0x0000FF00u) | ((s[2] >> 2) & 0x00FF0000u) | (s[3] & 0xFF000000u);
}

// textures.rs:309-362
void build_texture_mip(const sw_texel* prev, uint32_t pw, uint32_t ph, sw_texel* out)
{
	const uint32_t w = pw >> 1, h = ph >> 1;
	for(uint32_t y = 0; y < h; ++y)
		for(uint32_t x = 0; x < w; ++x)
		{
			const sw_texel p00 = prev[(y * 2) * pw + x * 2];
			const sw_texel p01 = prev[(y * 2 + 1) * pw + x * 2];
			const sw_texel p10 = prev[(y * 2) * pw + x * 2 + 1];
			const sw_texel p11 = prev[(y * 2 + 1) * pw + x * 2 + 1];
			sw_texel d;
			d.diffuse = average_4(p00.diffuse, p01.diffuse, p10.diffuse, p11.diffuse);
			V3 n[4];
			float r[4];
			unpack_normal_roughness(p00.packed_normal_roughness, n[0], r[0]);
			unpack_normal_roughness(p01.packed_normal_roughness, n[1], r[1]);
			unpack_normal_roughness(p10.packed_normal_roughness, n[2], r[2]);
			unpack_normal_roughness(p11.packed_normal_roughness, n[3], r[3]);
			const V3 sum = ((n[0] + n[1]) + n[2]) + n[3];
			const float len = std::fmax(length(sum), 0.000001f);
			const V3 dn = sum / len;
			const float half_cos = len * 0.25f;
			const float deviation = (1.0f - half_cos) / (1.0f - 0.5f);
			const float rs = ((r[0] + r[1]) + r[2]) + r[3];
			const float dr = std::fmin(std::fmax(std::fma(rs, 0.25f, deviation), MIN_VALID_ROUGHNESS), 1.0f);
			d.packed_normal_roughness = pack_normal_roughness(dn, dr);
			out[y * w + x] = d;
		}
}

uint64_t build_texture_mips(const sw_texel* mip0, uint32_t w, uint32_t h, sw_texel* out)
{
	const sw_texel* prev = mip0;
	uint32_t pw = w, ph = h;
	uint64_t written = 0;
	for(uint32_t mip = 1; mip <= MAX_MIP; ++mip)
	{
		build_texture_mip(prev, pw, ph, out + written);
		prev = out + written;
		written += uint64_t(pw >> 1) * (ph >> 1);
		pw >>= 1;
		ph >>= 1;
	}
	return written;
}

// textures.rs:385-415
void build_lite_mips(std::vector<uint32_t>& lite, uint64_t offset0, uint32_t w, uint32_t h)
{
	uint64_t prev = offset0;
	uint32_t pw = w, ph = h;
	for(uint32_t mip = 1; mip <= MAX_MIP; ++mip)
	{
		const uint32_t mw = pw >> 1, mh = ph >> 1;
		const uint64_t cur = lite.size();
		lite.resize(cur + uint64_t(mw) * mh);
		for(uint32_t y = 0; y < mh; ++y)
			for(uint32_t x = 0; x < mw; ++x)
				lite[cur + y * mw + x] = average_4(lite[prev + (y * 2) * pw + x * 2], lite[prev + (y * 2 + 1) * pw + x * 2],
					lite[prev + (y * 2) * pw + x * 2 + 1], lite[prev + (y * 2 + 1) * pw + x * 2 + 1]);
		prev = cur;
		pw = mw;
		ph = mh;
	}
}

} // namespace

// =============================================================================================
// Map generation
// =============================================================================================
namespace
{

const int DIRS[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

uint32_t build_kd(swh_map& m, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, uint32_t z0, uint32_t z1)
{
	if(x1 - x0 == 1 && y1 - y0 == 1 && z1 - z0 == 1)
		return FIRST_LEAF_INDEX + m.cell_leaf[cell_index(m, x0, y0, z0)];
	const uint32_t dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
	int axis = 0;
	if(dy > dx && dy >= dz)
		axis = 1;
	else if(dz > dx && dz > dy)
		axis = 2;
	uint32_t lo, hi;
	Node n;
	n.plane.vec = v3(axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f);
	if(axis == 0)
	{
		const uint32_t mid = x0 + dx / 2;
		lo = build_kd(m, x0, mid, y0, y1, z0, z1);
		hi = build_kd(m, mid, x1, y0, y1, z0, z1);
		n.plane.dist = CELL * float(mid);
	}
	else if(axis == 1)
	{
		const uint32_t mid = y0 + dy / 2;
		lo = build_kd(m, x0, x1, y0, mid, z0, z1);
		hi = build_kd(m, x0, x1, mid, y1, z0, z1);
		n.plane.dist = CELL * float(mid);
	}
	else
	{
		const uint32_t mid = z0 + dz / 2;
		lo = build_kd(m, x0, x1, y0, y1, z0, mid);
		hi = build_kd(m, x0, x1, y0, y1, mid, z1);
		n.plane.dist = CELL * float(mid);
	}
	// children[0] is the plane's front (positive) side, drawn last (partial_renderer.rs:2506-2514)
	n.children[0] = hi;
	n.children[1] = lo;
	m.nodes.push_back(n);
	return uint32_t(m.nodes.size() - 1);
}

// Quad of cell face `dir` (0..5) of cell (x, y, z).
void face_quad(uint32_t x, uint32_t y, uint32_t z, int dir, V3 out[4])
{
	const float x0 = CELL * float(x), x1 = x0 + CELL, y0 = CELL * float(y), y1 = y0 + CELL, z0 = CELL * float(z), z1 = z0 + CELL;
	switch(dir)
	{
	case 0: out[0] = v3(x0, y0, z0); out[1] = v3(x0, y1, z0); out[2] = v3(x0, y1, z1); out[3] = v3(x0, y0, z1); break;
	case 1: out[0] = v3(x1, y0, z0); out[1] = v3(x1, y1, z0); out[2] = v3(x1, y1, z1); out[3] = v3(x1, y0, z1); break;
	case 2: out[0] = v3(x0, y0, z0); out[1] = v3(x1, y0, z0); out[2] = v3(x1, y0, z1); out[3] = v3(x0, y0, z1); break;
	case 3: out[0] = v3(x0, y1, z0); out[1] = v3(x1, y1, z0); out[2] = v3(x1, y1, z1); out[3] = v3(x0, y1, z1); break;
	case 4: out[0] = v3(x0, y0, z0); out[1] = v3(x1, y0, z0); out[2] = v3(x1, y1, z0); out[3] = v3(x0, y1, z0); break;
	default: out[0] = v3(x0, y0, z1); out[1] = v3(x1, y0, z1); out[2] = v3(x1, y1, z1); out[3] = v3(x0, y1, z1); break;
	}
}

// Segment visibility through the grid (3D DDA over cells); solid cells block.
bool segment_clear(const swh_map& m, V3 a, V3 b)
{
	const V3 d = b - a;
	int c[3] = {int(std::floor(a.x / CELL)), int(std::floor(a.y / CELL)), int(std::floor(a.z / CELL))};
	const int end[3] = {int(std::floor(b.x / CELL)), int(std::floor(b.y / CELL)), int(std::floor(b.z / CELL))};
	const float dv[3] = {d.x, d.y, d.z};
	const float av[3] = {a.x, a.y, a.z};
	int step[3];
	float t_max[3], t_delta[3];
	for(int i = 0; i < 3; ++i)
	{
		step[i] = dv[i] > 0 ? 1 : (dv[i] < 0 ? -1 : 0);
		if(step[i] == 0)
		{
			t_max[i] = 1e30f;
			t_delta[i] = 1e30f;
		}
		else
		{
			const float boundary = CELL * float(c[i] + (step[i] > 0 ? 1 : 0));
			t_max[i] = (boundary - av[i]) / dv[i];
			t_delta[i] = CELL / std::fabs(dv[i]);
		}
	}
	for(int guard = 0; guard < 4096; ++guard)
	{
		if(c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= int(m.gx) || c[1] >= int(m.gy) || c[2] >= int(m.gz))
			return false;
		if(m.cell[cell_index(m, uint32_t(c[0]), uint32_t(c[1]), uint32_t(c[2]))] == 1)
			return false;
		if(c[0] == end[0] && c[1] == end[1] && c[2] == end[2])
			return true;
		int i = 0;
		if(t_max[1] < t_max[i])
			i = 1;
		if(t_max[2] < t_max[i])
			i = 2;
		if(t_max[i] > 1.0f)
			return true;
		c[i] += step[i];
		t_max[i] += t_delta[i];
	}
	return false;
}

// Sun visibility: walk towards the sun until a sky cell is reached.
bool sun_visible(const swh_map& m, V3 p, V3 to_sun)
{
	// march towards the sun until a sky cell (lit) or a solid cell / the grid border (shadowed)
	V3 q = p;
	for(int i = 0; i < 4096; ++i)
	{
		q = q + to_sun * 8.0f;
		const int cx = int(std::floor(q.x / CELL)), cy = int(std::floor(q.y / CELL)), cz = int(std::floor(q.z / CELL));
		if(cx < 0 || cy < 0 || cz < 0 || cx >= int(m.gx) || cy >= int(m.gy) || cz >= int(m.gz))
			return false;
		const uint8_t c = m.cell[cell_index(m, uint32_t(cx), uint32_t(cy), uint32_t(cz))];
		if(c == 2)
			return true;
		if(c == 1)
			return false;
	}
	return false;
}

void generate_material_texture(swh_map& m, uint32_t mat, uint32_t size, bool normal_map, float roughness, bool metal,
	bool alpha, uint32_t seed)
{
	Rng rng(seed * 7919u + mat);
	const float base[3] = {rng.uniform(0.25f, 1.0f), rng.uniform(0.25f, 1.0f), rng.uniform(0.25f, 1.0f)};
	const uint64_t off0 = m.texels.size();
	m.texels.resize(off0 + uint64_t(size) * size);
	std::vector<float> height;
	if(normal_map)
	{
		// Gaussian-bumped height field, sigma = 2 texels, splatted with wrap-around.
		height.assign(uint64_t(size) * size, 0.0f);
		const uint32_t bumps = size * size / 48;
		const float sigma = 2.0f;
		for(uint32_t b = 0; b < bumps; ++b)
		{
			const float cx = rng.uniform(0.0f, float(size)), cy = rng.uniform(0.0f, float(size));
			const float amp = rng.uniform(-1.0f, 1.0f);
			for(int dy = -6; dy <= 6; ++dy)
				for(int dx = -6; dx <= 6; ++dx)
				{
					const int px = int(std::floor(cx)) + dx, py = int(std::floor(cy)) + dy;
					const float fx = float(px) + 0.5f - cx, fy = float(py) + 0.5f - cy;
					const uint32_t wx = uint32_t((px % int(size) + int(size)) % int(size));
					const uint32_t wy = uint32_t((py % int(size) + int(size)) % int(size));
					height[wy * size + wx] += amp * std::exp(-(fx * fx + fy * fy) / (2.0f * sigma * sigma));
				}
		}
	}
	for(uint32_t y = 0; y < size; ++y)
		for(uint32_t x = 0; x < size; ++x)
		{
			const float n0 = fractal_noise(float(x), float(y), size, seed * 31u + mat * 3u + 0u);
			const float n1 = fractal_noise(float(x), float(y), size, seed * 31u + mat * 3u + 1u);
			const float k = 0.45f + 0.75f * n0;
			const uint32_t r = to_u8(255.0f * base[0] * k * (0.85f + 0.3f * n1));
			const uint32_t g = to_u8(255.0f * base[1] * k);
			const uint32_t b = to_u8(255.0f * base[2] * k * (1.15f - 0.3f * n1));
			const uint32_t a = alpha ? to_u8(255.0f * fractal_noise(float(x), float(y), size, seed * 31u + mat * 3u + 2u)) : 255u;
			uint32_t normal_rgb = 0x808080u | 0xFFu;
			if(normal_map)
			{
				auto h = [&](int xx, int yy) {
					return height[uint32_t((yy % int(size) + int(size)) % int(size)) * size +
						uint32_t((xx % int(size) + int(size)) % int(size))];
				};
				const float strength = 3.0f;
				const float dhx = (h(int(x) + 1, int(y)) - h(int(x) - 1, int(y))) * 0.5f * strength;
				const float dhy = (h(int(x), int(y) + 1) - h(int(x), int(y) - 1)) * 0.5f * strength;
				V3 n = v3(-dhx, -dhy, 1.0f);
				n = n / length(n);
				normal_rgb = (uint32_t(to_u8(128.0f + 127.0f * n.x)) << 16) | (uint32_t(to_u8(128.0f + 127.0f * n.y)) << 8) |
					uint32_t(to_u8(128.0f + 127.0f * n.z));
			}
			sw_texel& t = m.texels[off0 + uint64_t(y) * size + x];
			t.diffuse = (a << 24) | (r << 16) | (g << 8) | b;
			t.packed_normal_roughness = make_texel_normal_roughness(normal_rgb, normal_map, roughness);
		}
	uint32_t flags = 0;
	if(normal_map)
		flags |= SW_TEX_HAS_NORMAL_MAP;
	if(roughness < 1.0f)
		flags |= SW_TEX_HAS_NON_ONE_ROUGHNESS;
	if(metal)
		flags |= SW_TEX_IS_METAL;
	m.materials[mat].first_level = uint32_t(m.levels.size());
	m.levels.push_back(sw_texture_level{off0, size, size, flags, 0});
	// mips (textures.rs:288-304)
	std::vector<sw_texel> mips(uint64_t(size) * size / 3 + 16);
	const uint64_t n = build_texture_mips(m.texels.data() + off0, size, size, mips.data());
	uint64_t off = m.texels.size();
	m.texels.insert(m.texels.end(), mips.begin(), mips.begin() + n);
	uint32_t w = size;
	for(uint32_t mip = 1; mip <= MAX_MIP; ++mip)
	{
		w >>= 1;
		m.levels.push_back(sw_texture_level{off, w, w, flags, 0});
		off += uint64_t(w) * w;
	}
}

uint32_t add_lite_texture(swh_map& m, uint32_t size, uint32_t seed, bool round_alpha)
{
	const uint64_t off0 = m.lite.size();
	m.lite.resize(off0 + uint64_t(size) * size);
	Rng rng(seed);
	const float base[3] = {rng.uniform(0.2f, 1.0f), rng.uniform(0.2f, 1.0f), rng.uniform(0.2f, 1.0f)};
	for(uint32_t y = 0; y < size; ++y)
		for(uint32_t x = 0; x < size; ++x)
		{
			const float n = fractal_noise(float(x), float(y), size, seed);
			uint32_t a = 255;
			if(round_alpha)
			{
				const float dx = float(x) + 0.5f - float(size) * 0.5f, dy = float(y) + 0.5f - float(size) * 0.5f;
				const float r = std::sqrt(dx * dx + dy * dy) / (float(size) * 0.5f);
				a = to_u8(255.0f * std::min(1.0f, std::max(0.0f, (1.0f - r) * 3.0f)) * (0.6f + 0.4f * n));
			}
			m.lite[off0 + uint64_t(y) * size + x] = (a << 24) | (uint32_t(to_u8(255.0f * base[0] * (0.4f + 0.8f * n))) << 16) |
				(uint32_t(to_u8(255.0f * base[1] * (0.4f + 0.8f * n))) << 8) | uint32_t(to_u8(255.0f * base[2] * (0.4f + 0.8f * n)));
		}
	const uint32_t first = uint32_t(m.levels.size());
	m.levels.push_back(sw_texture_level{off0, size, size, SW_TEX_LITE_LEVEL, 0});
	build_lite_mips(m.lite, off0, size, size);
	uint64_t off = off0 + uint64_t(size) * size;
	uint32_t w = size;
	for(uint32_t mip = 1; mip <= MAX_MIP; ++mip)
	{
		w >>= 1;
		m.levels.push_back(sw_texture_level{off, w, w, SW_TEX_LITE_LEVEL, 0});
		off += uint64_t(w) * w;
	}
	return first;
}

// Skybox side textures (textures.rs:183-286): a sky gradient with noise clouds, Color32 and Color64.
void build_skybox(swh_map& m, uint32_t seed)
{
	const uint32_t size = 128;
	m.sky_side_size = size;
	for(uint32_t side = 0; side < 6; ++side)
	{
		uint32_t s = size;
		std::vector<uint32_t> mip0(uint64_t(size) * size);
		for(uint32_t y = 0; y < size; ++y)
			for(uint32_t x = 0; x < size; ++x)
			{
				const float n = fractal_noise(float(x), float(y), size, seed * 11u + side);
				const float up = side == 5 ? 1.0f : (side == 4 ? 0.0f : 1.0f - float(y) / float(size));
				const float cloud = std::max(0.0f, n - 0.45f) * 2.0f;
				const float r = 90.0f + 80.0f * up + 120.0f * cloud, g = 140.0f + 60.0f * up + 110.0f * cloud,
							b = 210.0f + 40.0f * up + 40.0f * cloud;
				mip0[y * size + x] = (uint32_t(to_u8(r)) << 16) | (uint32_t(to_u8(g)) << 8) | uint32_t(to_u8(b));
			}
		std::vector<uint32_t> cur = mip0;
		for(uint32_t mip = 0; mip <= MAX_MIP; ++mip)
		{
			for(uint32_t c : cur)
			{
				m.sky32.push_back(c);
				// 64-bit sides keep the 8-bit value in 16-bit lanes scaled by 1 (brightness 1.0)
				m.sky64.push_back(uint64_t(c & 0xFFu) | (uint64_t((c >> 8) & 0xFFu) << 16) | (uint64_t((c >> 16) & 0xFFu) << 32));
			}
			const uint32_t ns = s >> 1;
			std::vector<uint32_t> next(uint64_t(ns) * ns);
			for(uint32_t y = 0; y < ns; ++y)
				for(uint32_t x = 0; x < ns; ++x)
					next[y * ns + x] = average_4(cur[(2 * y) * s + 2 * x], cur[(2 * y + 1) * s + 2 * x], cur[(2 * y) * s + 2 * x + 1],
						cur[(2 * y + 1) * s + 2 * x + 1]);
			cur.swap(next);
			s = ns;
		}
	}
}

// Light grid samples at cell centres from the static lights (the lightmapper's light grid pass).
void build_light_grid(swh_map& m)
{
	const float CELL = 128.0f;
	m.light_grid.assign(m.cell.size(), LightGridElement{});
	for(uint32_t c : m.empty_cells)
	{
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		const V3 pos = v3(CELL * (float(x) + 0.5f), CELL * (float(y) + 0.5f), CELL * (float(z) + 0.5f));
		LightCubeAccum acc;
		for(const PointLight& l : m.lights)
		{
			const V3 d = l.pos - pos;
			const float d2 = std::max(dot(d, d), 64.0f * 64.0f);
			if(!segment_clear(m, pos, l.pos))
				continue;
			const float k = l.intensity / d2;
			const float light[3] = {l.color.x * k, l.color.y * k, l.color.z * k};
			acc.add_sample(d, light);
		}
		if(m.sun_intensity > 0.0f)
		{
			const V3 to_sun = m.sun_dir * -1.0f;
			if(sun_visible(m, pos, to_sun))
			{
				const float light[3] = {m.sun_intensity, m.sun_intensity * 0.95f, m.sun_intensity * 0.85f};
				acc.add_sample(to_sun, light);
			}
		}
		const float ambient[3] = {0.06f, 0.06f, 0.06f};
		acc.add_constant(ambient);
		m.light_grid[c] = acc.convert();
	}
}

// A geodesic sphere (icosahedron faces split into freq^2 triangles, 20 * 7^2 = 980 triangles) with
// `frames` frames of a seeded radial wobble: the procedural vertex-animated model.
void build_model_mesh(swh_map& m, uint32_t seed, uint32_t freq, uint32_t frames, float radius)
{
	const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
	const V3 ico[12] = {v3(-1, t, 0), v3(1, t, 0), v3(-1, -t, 0), v3(1, -t, 0), v3(0, -1, t), v3(0, 1, t), v3(0, -1, -t),
		v3(0, 1, -t), v3(t, 0, -1), v3(t, 0, 1), v3(-t, 0, -1), v3(-t, 0, 1)};
	const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4}, {11, 10, 2},
		{10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10},
		{8, 6, 7}, {9, 8, 1}};
	ModelMesh& mm = m.model_mesh;
	std::vector<V3> dirs;
	std::vector<uint32_t>& tris = mm.tris;
	tris.clear();
	auto vertex_id = [&](V3 p) -> uint32_t {
		const V3 d = p / length(p);
		for(uint32_t i = 0; i < dirs.size(); ++i)
		{
			const V3 e = dirs[i] - d;
			if(dot(e, e) < 1e-10f)
				return i;
		}
		dirs.push_back(d);
		return uint32_t(dirs.size() - 1);
	};
	for(const auto& f : faces)
	{
		const V3 a = ico[f[0]], b = ico[f[1]], c = ico[f[2]];
		std::vector<uint32_t> grid;
		// row i has freq - i + 1 points: a + (b - a) * j / freq + (c - a) * i / freq
		for(uint32_t i = 0; i <= freq; ++i)
			for(uint32_t j = 0; j + i <= freq; ++j)
				grid.push_back(vertex_id(a + (b - a) * (float(j) / float(freq)) + (c - a) * (float(i) / float(freq))));
		auto at = [&](uint32_t i, uint32_t j) {
			uint32_t idx = 0;
			for(uint32_t r = 0; r < i; ++r)
				idx += freq - r + 1;
			return grid[idx + j];
		};
		for(uint32_t i = 0; i < freq; ++i)
			for(uint32_t j = 0; j + i < freq; ++j)
			{
				tris.insert(tris.end(), {at(i, j), at(i, j + 1), at(i + 1, j)});
				if(j + i + 1 < freq)
					tris.insert(tris.end(), {at(i, j + 1), at(i + 1, j + 1), at(i + 1, j)});
			}
	}
	mm.num_vertices = uint32_t(dirs.size());
	mm.num_frames = frames;
	mm.tc.resize(size_t(mm.num_vertices) * 2);
	for(uint32_t v = 0; v < mm.num_vertices; ++v)
	{
		const V3 d = dirs[v];
		mm.tc[v * 2 + 0] = 0.5f + std::atan2(d.y, d.x) * (0.5f / 3.14159265f);
		mm.tc[v * 2 + 1] = std::acos(std::max(-1.0f, std::min(1.0f, d.z))) * (1.0f / 3.14159265f);
	}
	Rng rng(seed * 31u + 5u);
	const float ka = rng.uniform(2.0f, 4.0f), kb = rng.uniform(2.0f, 4.0f), kc = rng.uniform(1.0f, 3.0f);
	mm.pos.resize(size_t(frames) * mm.num_vertices);
	mm.normal.assign(size_t(frames) * mm.num_vertices, v3(0, 0, 0));
	mm.frame_bbox.resize(frames);
	for(uint32_t f = 0; f < frames; ++f)
	{
		const float phase = 6.2831853f * float(f) / float(frames);
		V3* pos = &mm.pos[size_t(f) * mm.num_vertices];
		V3* nrm = &mm.normal[size_t(f) * mm.num_vertices];
		for(uint32_t v = 0; v < mm.num_vertices; ++v)
		{
			const V3 d = dirs[v];
			const float wob = 1.0f + 0.18f * std::sin(ka * d.x + kb * d.y + phase) * std::cos(kc * d.z - phase);
			pos[v] = d * (radius * wob);
		}
		for(size_t k = 0; k < tris.size(); k += 3)
		{
			const V3 a = pos[tris[k]], b = pos[tris[k + 1]], c = pos[tris[k + 2]];
			const V3 n = cross(b - a, c - a);
			nrm[tris[k]] = nrm[tris[k]] + n;
			nrm[tris[k + 1]] = nrm[tris[k + 1]] + n;
			nrm[tris[k + 2]] = nrm[tris[k + 2]] + n;
		}
		BBox bb{pos[0], pos[0]};
		for(uint32_t v = 0; v < mm.num_vertices; ++v)
		{
			nrm[v] = nrm[v] / std::max(length(nrm[v]), 1e-20f);
			bb.min = v3(std::min(bb.min.x, pos[v].x), std::min(bb.min.y, pos[v].y), std::min(bb.min.z, pos[v].z));
			bb.max = v3(std::max(bb.max.x, pos[v].x), std::max(bb.max.y, pos[v].y), std::max(bb.max.z, pos[v].z));
		}
		mm.frame_bbox[f] = bb;
	}
}

// Inline models: seeded doors (thin upright slabs) and platforms (flat slabs) standing in empty
// cells on a solid floor, each a box of six outward-facing quads after the leafs' polygons, ordered
// by a chain of submodel BSP nodes (one face per node; the faces of a convex box never hide one
// another, so any back-to-front walk of the chain is exact). The lightmapper lights them in their
// initial position, as the reference's lightmapper does.
void generate_submodels(swh_map& m, uint32_t count, uint32_t seed, uint32_t nmat, uint32_t texture_size)
{
	if(count == 0)
		return;
	Rng rng(uint64_t(seed) * 0xD1B54A32D192ED03ull + 77);
	std::vector<uint32_t> candidates;
	for(uint32_t c : m.empty_cells)
	{
		const uint32_t z = c / (m.gx * m.gy);
		if(z > 0 && m.cell[c - m.gx * m.gy] == 1)
			candidates.push_back(c);
	}
	for(uint32_t k = 0; k < count && !candidates.empty(); ++k)
	{
		const uint32_t pick = rng.below(uint32_t(candidates.size()));
		const uint32_t c = candidates[pick];
		candidates.erase(candidates.begin() + pick);
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		const bool door = k % 2 == 0;
		const bool along_x = rng.below(2) == 0;
		const V3 half = door ? (along_x ? v3(40.0f, 8.0f, 44.0f) : v3(8.0f, 40.0f, 44.0f)) : v3(32.0f, 32.0f, 8.0f);
		const V3 centre = v3(CELL * (float(x) + 0.5f), CELL * (float(y) + 0.5f), CELL * float(z) + (door ? 46.0f : 40.0f));
		const BBox b{centre - half, centre + half};
		Submodel sm;
		sm.first_polygon = uint32_t(m.polygons.size());
		sm.num_polygons = 6;
		sm.root_node = uint32_t(m.submodel_nodes.size());
		const uint32_t texture = rng.below(nmat);
		for(int dir = 0; dir < 6; ++dir)
		{
			const int axis = dir / 2;
			const bool positive = (dir & 1) != 0;
			const V3 normal = v3(axis == 0 ? (positive ? 1.0f : -1.0f) : 0.0f, axis == 1 ? (positive ? 1.0f : -1.0f) : 0.0f,
				axis == 2 ? (positive ? 1.0f : -1.0f) : 0.0f);
			// the face's four corners: the other two axes at their min / max
			const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
			V3 q[4];
			for(int k4 = 0; k4 < 4; ++k4)
			{
				float p[3];
				const float mn[3] = {b.min.x, b.min.y, b.min.z}, mx[3] = {b.max.x, b.max.y, b.max.z};
				p[axis] = positive ? mx[axis] : mn[axis];
				p[a1] = (k4 == 1 || k4 == 2) ? mx[a1] : mn[a1];
				p[a2] = (k4 >= 2) ? mx[a2] : mn[a2];
				q[k4] = v3(p[0], p[1], p[2]);
			}
			// clockwise seen from the front: cross(e1, e2) points away from the normal
			if(dot(cross(q[1] - q[0], q[2] - q[1]), normal) > 0.0f)
				std::swap(q[1], q[3]);
			Polygon p;
			p.first_vertex = uint32_t(m.vertices.size());
			p.num_vertices = 4;
			for(int k4 = 0; k4 < 4; ++k4)
				m.vertices.push_back(q[k4]);
			p.plane = Plane{normal, dot(normal, q[0])};
			const float tshift0 = float(rng.below(texture_size)), tshift1 = float(rng.below(texture_size));
			if(axis == 0)
			{
				p.tex_coord_equation[0] = Plane{v3(0, 1, 0), tshift0};
				p.tex_coord_equation[1] = Plane{v3(0, 0, -1), tshift1};
			}
			else if(axis == 1)
			{
				p.tex_coord_equation[0] = Plane{v3(1, 0, 0), tshift0};
				p.tex_coord_equation[1] = Plane{v3(0, 0, -1), tshift1};
			}
			else
			{
				p.tex_coord_equation[0] = Plane{v3(1, 0, 0), tshift0};
				p.tex_coord_equation[1] = Plane{v3(0, -1, 0), tshift1};
			}
			for(int t = 0; t < 2; ++t)
			{
				float tmin = 1e30f, tmax = -1e30f;
				for(int k4 = 0; k4 < 4; ++k4)
				{
					const float tc = dot(p.tex_coord_equation[t].vec, q[k4]) + p.tex_coord_equation[t].dist;
					tmin = std::min(tmin, tc);
					tmax = std::max(tmax, tc);
				}
				p.tex_coord_min[t] = int32_t(std::floor(tmin));
				p.tex_coord_max[t] = int32_t(std::ceil(tmax));
			}
			p.lightmap_data_offset = 0;
			p.texture = texture;
			const uint32_t pi = uint32_t(m.polygons.size());
			m.polygons.push_back(p);
			m.polygon_leaf.push_back(NO_POLYGON_LEAF);
			SubmodelNode node;
			node.plane = p.plane;
			node.first_polygon = pi;
			node.num_polygons = 1;
			node.children[0] = dir + 1 < 6 ? uint32_t(m.submodel_nodes.size()) + 1u : 0xFFFFFFFFu;
			node.children[1] = 0xFFFFFFFFu;
			m.submodel_nodes.push_back(node);
		}
		m.submodels.push_back(sm);
		m.submodel_cell.push_back(c);
	}
}

// View portals on walls: in empty cells with a solid neighbour across x or y, a 64 x 80 quad one unit
// in front of the wall, facing into the cell, texture coordinates one per unit along the wall and
// down. Mirrors, camera portals (a camera in another empty cell, looking in a seeded direction) and
// parallax portals (the view point moved by the offset to another empty cell) alternate.
void generate_view_portals(swh_map& m, uint32_t count, uint32_t seed)
{
	if(count == 0)
		return;
	Rng rng(uint64_t(seed) * 0x94D049BB133111EBull + 91);
	std::vector<std::pair<uint32_t, int>> walls;
	for(uint32_t c : m.empty_cells)
	{
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		for(int dir = 0; dir < 4; ++dir)
		{
			const int nx = int(x) + DIRS[dir][0], ny = int(y) + DIRS[dir][1];
			if(m.cell[cell_index(m, uint32_t(nx), uint32_t(ny), z)] == 1)
				walls.push_back({c, dir});
		}
	}
	const std::vector<std::pair<uint32_t, int>> all_walls = walls;
	for(uint32_t k = 0; k < count && !walls.empty(); ++k)
	{
		const uint32_t pick = rng.below(uint32_t(walls.size()));
		const uint32_t c = walls[pick].first;
		const int dir = walls[pick].second;
		walls.erase(walls.begin() + pick);
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		ViewPortalDef p{};
		p.kind = k % 3;
		p.blending_mode = SW_BLEND_NONE;
		const V3 normal = v3(-float(DIRS[dir][0]), -float(DIRS[dir][1]), 0.0f); // into the cell
		const V3 centre = v3(CELL * (float(x) + 0.5f), CELL * (float(y) + 0.5f), CELL * float(z) + 64.0f);
		const V3 on_wall = centre - normal * (CELL * 0.5f - 1.0f);
		// the screen's right for a viewer facing the wall (cameras map world +z to screen up)
		const V3 right = cross(normal * -1.0f, v3(0.0f, 0.0f, 1.0f));
		V3 q[4] = {on_wall - right * 32.0f - v3(0, 0, 40.0f), on_wall + right * 32.0f - v3(0, 0, 40.0f),
			on_wall + right * 32.0f + v3(0, 0, 40.0f), on_wall - right * 32.0f + v3(0, 0, 40.0f)};
		// clockwise on screen seen from the front (the rasterizer's order): cross(e1, e2) points away
		// from the normal
		if(dot(cross(q[1] - q[0], q[2] - q[1]), normal) > 0.0f)
			std::swap(q[1], q[3]);
		for(int i = 0; i < 4; ++i)
			p.vertices[i] = q[i];
		p.plane = Plane{normal, dot(normal, q[0])};
		// Two texels per unit, (0, 0) at the top-left corner seen from the front. A mirror's view
		// matrix maps (u, v, distance along the normal) to the screen, so with u to the right the
		// mirrored scene comes out counter-clockwise and nothing is rasterized: mirrors take u to
		// the left, which shows the same reflection. Camera and parallax views keep u to the right.
		const V3 u = p.kind == PORTAL_MIRROR ? right * -2.0f : right * 2.0f;
		p.tex_coord_equation[0] = Plane{u, -dot(u, on_wall) + 64.0f};
		p.tex_coord_equation[1] = Plane{v3(0, 0, -2.0f), 2.0f * (on_wall.z + 40.0f)};
		const uint32_t other = m.empty_cells[rng.below(uint32_t(m.empty_cells.size()))];
		const uint32_t ox = other % m.gx, oy = (other / m.gx) % m.gy, oz = other / (m.gx * m.gy);
		p.camera_position = v3(CELL * (float(ox) + 0.5f), CELL * (float(oy) + 0.5f), CELL * float(oz) + 64.0f);
		p.camera_yaw = rng.uniform(0.0f, 6.2831853f);
		p.camera_pitch = rng.uniform(-0.3f, 0.3f);
		p.camera_fov = 1.5707963f;
		// parallax: the view from the same spot in front of another wall facing the same way
		uint32_t target = c;
		const uint32_t first = rng.below(uint32_t(all_walls.size()));
		for(size_t i = 0; i < all_walls.size(); ++i)
		{
			const auto& w = all_walls[(first + i) % all_walls.size()];
			if(w.second == dir && w.first != c)
			{
				target = w.first;
				break;
			}
		}
		const uint32_t tx = target % m.gx, ty = (target / m.gx) % m.gy, tz = target / (m.gx * m.gy);
		p.transform =
			from_translation(v3(CELL * (float(tx) - float(x)), CELL * (float(ty) - float(y)), CELL * (float(tz) - float(z))));
		// every second portal carries a texture overlay (get_portal_texture, test_game/world_spawn.rs:1009-1041:
		// light scale 1, no added light): an opaque one averaged in (the default blending), or a
		// round alpha-masked one blended by its alpha
		if(k % 4 == 1 && !m.model_levels.empty())
		{
			p.texture_level = m.model_levels[(k / 4) % m.model_levels.size()];
			p.texture_blending_mode = SW_BLEND_AVERAGE;
		}
		else if(k % 4 == 3 && !m.decal_levels.empty())
		{
			p.texture_level = m.decal_levels[(k / 4) % m.decal_levels.size()];
			p.texture_blending_mode = SW_BLEND_ALPHA_BLEND;
		}
		m.view_portals.push_back(p);
	}
}

void light_polygons(swh_map& m)
{
	const bool directional = m.config.directional_lightmaps != 0;
	m.lm_simple.assign(3, 0.0f); // element 0 is reserved: offset 0 means "no lightmap"
	m.lm_dir.assign(1, sw_directional_lightmap_element{});
	for(uint32_t pi = 0; pi < m.polygons.size(); ++pi)
	{
		Polygon& p = m.polygons[pi];
		if(m.materials[p.texture].sky)
		{
			p.lightmap_data_offset = 0;
			continue;
		}
		const uint32_t lw = uint32_t(((p.tex_coord_max[0] + int32_t(LIGHTMAP_SCALE) - 1) >> LIGHTMAP_SCALE_LOG2) -
			(p.tex_coord_min[0] >> LIGHTMAP_SCALE_LOG2) + 1);
		const uint32_t lh = uint32_t(((p.tex_coord_max[1] + int32_t(LIGHTMAP_SCALE) - 1) >> LIGHTMAP_SCALE_LOG2) -
			(p.tex_coord_min[1] >> LIGHTMAP_SCALE_LOG2) + 1);
		// lightmap basis (common/lightmap.rs:36-63)
		const float s = 1.0f / float(LIGHTMAP_SCALE);
		M4 b;
		b.c[0] = v4(p.tex_coord_equation[0].vec.x * s, p.tex_coord_equation[0].vec.y * s, p.tex_coord_equation[0].vec.z * s,
			p.tex_coord_equation[0].dist * s);
		b.c[1] = v4(p.tex_coord_equation[1].vec.x * s, p.tex_coord_equation[1].vec.y * s, p.tex_coord_equation[1].vec.z * s,
			p.tex_coord_equation[1].dist * s);
		b.c[2] = extend(p.plane.vec, -p.plane.dist);
		b.c[3] = v4(0, 0, 0, 1);
		M4 inv;
		if(!invert(transpose(b), inv))
		{
			p.lightmap_data_offset = 0;
			continue;
		}
		const V3 u_vec = truncate(inv.c[0]), v_vec = truncate(inv.c[1]);
		const V3 start = truncate(inv.c[3]) + u_vec * float(p.tex_coord_min[0] >> LIGHTMAP_SCALE_LOG2) +
			v_vec * float(p.tex_coord_min[1] >> LIGHTMAP_SCALE_LOG2);
		const V3 n = p.plane.vec / length(p.plane.vec);
		const V3 un = u_vec / length(u_vec), vn = v_vec / length(v_vec);
		// polygon extent, to keep samples inside the face (no light leaking through edges)
		V3 pmin = m.vertices[p.first_vertex], pmax = pmin;
		for(uint32_t k = 1; k < p.num_vertices; ++k)
		{
			const V3 q = m.vertices[p.first_vertex + k];
			pmin = v3(std::min(pmin.x, q.x), std::min(pmin.y, q.y), std::min(pmin.z, q.z));
			pmax = v3(std::max(pmax.x, q.x), std::max(pmax.y, q.y), std::max(pmax.z, q.z));
		}
		p.lightmap_data_offset = uint32_t(m.lm_dir.size());
		for(uint32_t j = 0; j < lh; ++j)
			for(uint32_t i = 0; i < lw; ++i)
			{
				V3 pos = start + u_vec * float(i) + v_vec * float(j);
				pos = v3(std::min(std::max(pos.x, pmin.x + 1.0f), pmax.x - 1.0f), std::min(std::max(pos.y, pmin.y + 1.0f), pmax.y - 1.0f),
					std::min(std::max(pos.z, pmin.z + 1.0f), pmax.z - 1.0f));
				pos = pos + n * 2.0f;
				float total[3] = {0, 0, 0};
				V3 dir_sum = v3(0, 0, 0);
				float weight = 0.0f;
				float col_sum[3] = {0, 0, 0};
				auto add_light = [&](V3 to_light_n, float cos_a, float intensity, V3 color) {
					const float li = intensity * cos_a;
					total[0] += color.x * li;
					total[1] += color.y * li;
					total[2] += color.z * li;
					const V3 ts = v3(dot(to_light_n, un), dot(to_light_n, vn), dot(to_light_n, n));
					dir_sum = dir_sum + ts * intensity;
					weight += intensity;
					col_sum[0] += color.x * intensity;
					col_sum[1] += color.y * intensity;
					col_sum[2] += color.z * intensity;
				};
				for(const PointLight& l : m.lights)
				{
					const V3 d = l.pos - pos;
					const float d2 = std::max(dot(d, d), 64.0f * 64.0f);
					const float cos_a = dot(n, d) / std::sqrt(dot(d, d));
					if(cos_a <= 0.0f || !segment_clear(m, pos, l.pos))
						continue;
					add_light(d / std::sqrt(dot(d, d)), cos_a, l.intensity / d2, l.color);
				}
				if(m.sun_intensity > 0.0f)
				{
					const V3 to_sun = m.sun_dir * -1.0f;
					const float cos_a = dot(n, to_sun);
					if(cos_a > 0.0f && sun_visible(m, pos, to_sun))
						add_light(to_sun, cos_a, m.sun_intensity, v3(1.0f, 0.95f, 0.85f));
				}
				const float ambient = 0.06f;
				float* simple = nullptr;
				m.lm_simple.resize(m.lm_simple.size() + 3);
				simple = &m.lm_simple[m.lm_simple.size() - 3];
				simple[0] = total[0] + ambient;
				simple[1] = total[1] + ambient;
				simple[2] = total[2] + ambient;
				sw_directional_lightmap_element e{};
				const float dir_len = length(dir_sum);
				const float deviation = weight > 0.0f ? std::min(1.0f, std::max(0.0f, 1.0f - dir_len / weight)) : 1.0f;
				// constant part: ambient plus the share of light that does not follow the main direction
				for(int c = 0; c < 3; ++c)
					e.ambient_light[c] = ambient + total[c] * 0.35f * deviation;
				if(weight > 0.0f && dir_len > 0.0f)
				{
					const V3 dn = dir_sum / dir_len;
					// directional part carries the remaining intensity along the dominant direction
					const float scale = (1.0f - 0.35f * deviation) * (total[0] + total[1] + total[2]) / 3.0f;
					e.light_direction_vector_scaled[0] = dn.x * scale;
					e.light_direction_vector_scaled[1] = dn.y * scale;
					e.light_direction_vector_scaled[2] = dn.z * scale;
					const float cmax = std::max(col_sum[0], std::max(col_sum[1], col_sum[2]));
					for(int c = 0; c < 3; ++c)
						e.directional_light_color[c] = cmax > 0.0f ? col_sum[c] / cmax : 0.0f;
				}
				e.directional_light_deviation = deviation;
				m.lm_dir.push_back(e);
			}
		if(!directional)
			continue;
	}
	if(!directional)
		m.lm_dir.clear();
}

} // namespace

// =============================================================================================
// C API
// =============================================================================================
extern "C" {

int32_t swh_pack_normal_roughness_from_rgb(uint32_t normal_rgb, float roughness)
{
	return make_texel_normal_roughness(normal_rgb, true, roughness);
}

uint64_t swh_build_texture_mips(const sw_texel* mip0, uint32_t w, uint32_t h, sw_texel* out)
{
	return build_texture_mips(mip0, w, h, out);
}

int swh_map_create(const swh_map_config* cfg, swh_map** out)
{
	if(!cfg || !out || cfg->grid[0] < 3 || cfg->grid[1] < 3 || cfg->grid[2] < 3 || cfg->grid[0] * cfg->grid[1] * cfg->grid[2] > (1u << 22) ||
		cfg->texture_size < 16 || (cfg->texture_size & (cfg->texture_size - 1)) || cfg->num_materials == 0)
		return -1;
	swh_map* m = new(std::nothrow) swh_map();
	if(!m)
		return -2;
	m->config = *cfg;
	m->gx = cfg->grid[0];
	m->gy = cfg->grid[1];
	m->gz = cfg->grid[2];
	const uint32_t ncell = m->gx * m->gy * m->gz;
	Rng rng(uint64_t(cfg->seed) * 0x9E3779B97F4A7C15ull + 1);

	// cells: perimeter solid, interior solid with probability p, then keep one connected region
	m->cell.assign(ncell, 1);
	for(uint32_t z = 1; z + 1 < m->gz; ++z)
		for(uint32_t y = 1; y + 1 < m->gy; ++y)
			for(uint32_t x = 1; x + 1 < m->gx; ++x)
				m->cell[cell_index(*m, x, y, z)] = rng.uniform(0.0f, 1.0f) < cfg->solid_probability ? 1 : 0;
	std::vector<uint32_t> empties;
	for(uint32_t i = 0; i < ncell; ++i)
		if(m->cell[i] == 0)
			empties.push_back(i);
	if(empties.empty())
	{
		const uint32_t c = cell_index(*m, m->gx / 2, m->gy / 2, m->gz / 2);
		m->cell[c] = 0;
		empties.push_back(c);
	}
	{
		const uint32_t start = empties[rng.below(uint32_t(empties.size()))];
		std::vector<uint8_t> seen(ncell, 0);
		std::vector<uint32_t> stack{start};
		seen[start] = 1;
		while(!stack.empty())
		{
			const uint32_t c = stack.back();
			stack.pop_back();
			const uint32_t x = c % m->gx, y = (c / m->gx) % m->gy, z = c / (m->gx * m->gy);
			for(const auto& d : DIRS)
			{
				const int nx = int(x) + d[0], ny = int(y) + d[1], nz = int(z) + d[2];
				if(nx < 0 || ny < 0 || nz < 0 || nx >= int(m->gx) || ny >= int(m->gy) || nz >= int(m->gz))
					continue;
				const uint32_t n = cell_index(*m, uint32_t(nx), uint32_t(ny), uint32_t(nz));
				if(!seen[n] && m->cell[n] == 0)
				{
					seen[n] = 1;
					stack.push_back(n);
				}
			}
		}
		for(uint32_t i = 0; i < ncell; ++i)
			if(m->cell[i] == 0 && !seen[i])
				m->cell[i] = 1;
	}
	for(uint32_t i = 0; i < ncell; ++i)
		if(m->cell[i] == 0)
			m->empty_cells.push_back(i);
	// sky openings: top-perimeter cells above empty cells of the top interior layer
	if(cfg->sun)
		for(uint32_t y = 1; y + 1 < m->gy; ++y)
			for(uint32_t x = 1; x + 1 < m->gx; ++x)
				if(m->cell[cell_index(*m, x, y, m->gz - 2)] == 0 && rng.uniform(0.0f, 1.0f) < cfg->sky_fraction)
					m->cell[cell_index(*m, x, y, m->gz - 1)] = 2;

	// materials: the last one is the sky
	const uint32_t nmat = cfg->num_materials;
	m->materials.resize(nmat + 1);
	for(uint32_t i = 0; i < nmat; ++i)
	{
		Material& mt = m->materials[i];
		mt.sky = false;
		mt.roughness = 1.0f;
		mt.is_metal = false;
		mt.blending_mode = SW_BLEND_NONE;
		const float r = rng.uniform(0.0f, 1.0f);
		if(r < cfg->specular_fraction)
		{
			mt.roughness = rng.uniform(0.1f, 0.6f);
			mt.is_metal = rng.uniform(0.0f, 1.0f) < 0.5f;
		}
		if(rng.uniform(0.0f, 1.0f) < cfg->alpha_blend_fraction)
			mt.blending_mode = SW_BLEND_ALPHA_BLEND;
		generate_material_texture(*m, i, cfg->texture_size, cfg->normal_maps != 0, mt.roughness, mt.is_metal,
			mt.blending_mode == SW_BLEND_ALPHA_BLEND, cfg->seed);
	}
	m->materials[nmat] = Material{0, 1.0f, false, SW_BLEND_NONE, true};
	for(uint32_t i = 0; i < cfg->num_decal_textures; ++i)
		m->decal_levels.push_back(add_lite_texture(*m, 32, cfg->seed * 131u + 7u * i + 1u, true));
	for(uint32_t i = 0; i < cfg->num_model_textures; ++i)
		m->model_levels.push_back(add_lite_texture(*m, 64, cfg->seed * 137u + 11u * i + 3u, false));
	build_skybox(*m, cfg->seed);

	// leafs (one per cell) and their polygons
	m->cell_leaf.resize(ncell);
	m->leafs.resize(ncell);
	m->leaf_bbox.resize(ncell);
	for(uint32_t i = 0; i < ncell; ++i)
	{
		m->cell_leaf[i] = i;
		const uint32_t x = i % m->gx, y = (i / m->gx) % m->gy, z = i / (m->gx * m->gy);
		m->leaf_bbox[i] = BBox{v3(CELL * x, CELL * y, CELL * z), v3(CELL * (x + 1), CELL * (y + 1), CELL * (z + 1))};
	}
	for(uint32_t i = 0; i < ncell; ++i)
	{
		Leaf& leaf = m->leafs[i];
		leaf.first_polygon = uint32_t(m->polygons.size());
		leaf.num_polygons = 0;
		if(m->cell[i] != 0)
			continue;
		const uint32_t x = i % m->gx, y = (i / m->gx) % m->gy, z = i / (m->gx * m->gy);
		for(int dir = 0; dir < 6; ++dir)
		{
			const int nx = int(x) + DIRS[dir][0], ny = int(y) + DIRS[dir][1], nz = int(z) + DIRS[dir][2];
			const uint8_t nc = m->cell[cell_index(*m, uint32_t(nx), uint32_t(ny), uint32_t(nz))];
			if(nc == 0)
				continue;
			V3 q[4];
			face_quad(x, y, z, dir, q);
			const V3 normal = v3(-float(DIRS[dir][0]), -float(DIRS[dir][1]), -float(DIRS[dir][2]));
			// clockwise seen from the front: cross(e1, e2) points away from the normal
			if(dot(cross(q[1] - q[0], q[2] - q[1]), normal) > 0.0f)
				std::swap(q[1], q[3]);
			Polygon p;
			p.first_vertex = uint32_t(m->vertices.size());
			p.num_vertices = 4;
			for(int k = 0; k < 4; ++k)
				m->vertices.push_back(q[k]);
			p.plane.vec = normal;
			p.plane.dist = dot(normal, q[0]);
			const int axis = dir / 2;
			const float tshift0 = float(rng.below(cfg->texture_size)), tshift1 = float(rng.below(cfg->texture_size));
			if(axis == 0)
			{
				p.tex_coord_equation[0] = Plane{v3(0, 1, 0), tshift0};
				p.tex_coord_equation[1] = Plane{v3(0, 0, -1), tshift1};
			}
			else if(axis == 1)
			{
				p.tex_coord_equation[0] = Plane{v3(1, 0, 0), tshift0};
				p.tex_coord_equation[1] = Plane{v3(0, 0, -1), tshift1};
			}
			else
			{
				p.tex_coord_equation[0] = Plane{v3(1, 0, 0), tshift0};
				p.tex_coord_equation[1] = Plane{v3(0, -1, 0), tshift1};
			}
			for(int t = 0; t < 2; ++t)
			{
				float tmin = 1e30f, tmax = -1e30f;
				for(int k = 0; k < 4; ++k)
				{
					const float tc = dot(p.tex_coord_equation[t].vec, q[k]) + p.tex_coord_equation[t].dist;
					tmin = std::min(tmin, tc);
					tmax = std::max(tmax, tc);
				}
				p.tex_coord_min[t] = int32_t(std::floor(tmin));
				p.tex_coord_max[t] = int32_t(std::ceil(tmax));
			}
			p.lightmap_data_offset = 0;
			// walls/floor/ceiling draw from different material subsets
			p.texture = nc == 2 ? nmat : (hash32(i * 6u + uint32_t(dir) + cfg->seed * 977u) % nmat);
			m->polygons.push_back(p);
			m->polygon_leaf.push_back(i);
			leaf.num_polygons++;
		}
	}

	// portals between adjacent empty cells; the plane faces leafs[0]
	std::vector<std::vector<uint32_t>> leaf_portals(ncell);
	for(uint32_t i = 0; i < ncell; ++i)
	{
		if(m->cell[i] != 0)
			continue;
		const uint32_t x = i % m->gx, y = (i / m->gx) % m->gy, z = i / (m->gx * m->gy);
		for(int dir = 1; dir < 6; dir += 2)
		{
			const uint32_t nx = x + uint32_t(DIRS[dir][0]), ny = y + uint32_t(DIRS[dir][1]), nz = z + uint32_t(DIRS[dir][2]);
			const uint32_t n = cell_index(*m, nx, ny, nz);
			if(m->cell[n] != 0)
				continue;
			V3 q[4];
			face_quad(x, y, z, dir, q);
			Portal p;
			p.leafs[0] = i;
			p.leafs[1] = n;
			const V3 normal = v3(-float(DIRS[dir][0]), -float(DIRS[dir][1]), -float(DIRS[dir][2]));
			p.plane = Plane{normal, dot(normal, q[0])};
			p.first_vertex = uint32_t(m->vertices.size());
			p.num_vertices = 4;
			for(int k = 0; k < 4; ++k)
				m->vertices.push_back(q[k]);
			leaf_portals[i].push_back(uint32_t(m->portals.size()));
			leaf_portals[n].push_back(uint32_t(m->portals.size()));
			m->portals.push_back(p);
		}
	}
	for(uint32_t i = 0; i < ncell; ++i)
	{
		m->leafs[i].first_leaf_portal = uint32_t(m->leafs_portals.size());
		m->leafs[i].num_leaf_portals = uint32_t(leaf_portals[i].size());
		m->leafs_portals.insert(m->leafs_portals.end(), leaf_portals[i].begin(), leaf_portals[i].end());
	}
	build_kd(*m, 0, m->gx, 0, m->gy, 0, m->gz);

	// lights
	for(uint32_t i = 0; i < cfg->num_lights; ++i)
	{
		const uint32_t c = m->empty_cells[rng.below(uint32_t(m->empty_cells.size()))];
		const uint32_t x = c % m->gx, y = (c / m->gx) % m->gy, z = c / (m->gx * m->gy);
		PointLight l;
		l.pos = v3(CELL * (float(x) + rng.uniform(0.2f, 0.8f)), CELL * (float(y) + rng.uniform(0.2f, 0.8f)),
			CELL * (float(z) + rng.uniform(0.4f, 0.9f)));
		l.intensity = 9000.0f * rng.uniform(cfg->light_intensity_min, cfg->light_intensity_max);
		l.color = v3(1.0f, 1.0f, 1.0f);
		m->lights.push_back(l);
	}
	if(cfg->sun)
	{
		m->sun_dir = v3(0.35f, 0.25f, -1.0f) / length(v3(0.35f, 0.25f, -1.0f));
		m->sun_intensity = 0.8f;
	}
	generate_submodels(*m, cfg->num_submodels, cfg->seed, nmat, cfg->texture_size);
	generate_view_portals(*m, cfg->num_view_portals, cfg->seed);
	light_polygons(*m);
	build_light_grid(*m);
	build_model_mesh(*m, cfg->seed, 7, 8, 28.0f);
	{
		ModelMesh& mm = m->model_mesh;
		mm.abi_frame_vertices.resize(size_t(mm.num_frames) * mm.num_vertices * 6);
		for(size_t i = 0; i < size_t(mm.num_frames) * mm.num_vertices; ++i)
		{
			float* o = &mm.abi_frame_vertices[i * 6];
			o[0] = mm.pos[i].x, o[1] = mm.pos[i].y, o[2] = mm.pos[i].z;
			o[3] = mm.normal[i].x, o[4] = mm.normal[i].y, o[5] = mm.normal[i].z;
		}
		// wound outward counter-clockwise: swap the last two, as the front-end's own preparation does
		mm.abi_triangles.resize(mm.tris.size());
		for(size_t k = 0; k + 2 < mm.tris.size(); k += 3)
		{
			mm.abi_triangles[k] = uint16_t(mm.tris[k]);
			mm.abi_triangles[k + 1] = uint16_t(mm.tris[k + 2]);
			mm.abi_triangles[k + 2] = uint16_t(mm.tris[k + 1]);
		}
	}
	*out = m;
	return 0;
}

void swh_map_destroy(swh_map* m) { delete m; }

int swh_map_desc(const swh_map* m, sw_map_desc* d)
{
	if(!m || !d)
		return -1;
	std::memset(d, 0, sizeof(*d));
	d->texels = m->texels.data();
	d->num_texels = m->texels.size();
	d->lite_texels = m->lite.data();
	d->num_lite_texels = m->lite.size();
	d->lightmaps_simple = m->lm_simple.data();
	d->num_lightmap_elements = m->lm_simple.size() / 3;
	d->lightmaps_directional = m->lm_dir.data();
	d->num_directional_lightmap_elements = m->lm_dir.size();
	d->levels = m->levels.data();
	d->num_levels = uint32_t(m->levels.size());
	d->skybox_texels32 = m->sky32.data();
	d->num_skybox_texels32 = m->sky32.size();
	d->skybox_texels64 = m->sky64.data();
	d->num_skybox_texels64 = m->sky64.size();
	return 0;
}

int swh_map_info_get(const swh_map* m, swh_map_info* info)
{
	if(!m || !info)
		return -1;
	info->num_polygons = uint32_t(m->polygons.size());
	info->num_leafs = uint32_t(m->leafs.size());
	info->num_nodes = uint32_t(m->nodes.size());
	info->num_portals = uint32_t(m->portals.size());
	info->num_empty_cells = uint32_t(m->empty_cells.size());
	info->num_texels = m->texels.size();
	info->num_lightmap_elements = m->lm_simple.size() / 3;
	return 0;
}

int swh_map_view_portals(const swh_map* m, swh_view_portal* out, uint32_t capacity)
{
	if(!m)
		return -1;
	const uint32_t n = uint32_t(m->view_portals.size());
	if(!out)
		return int(n);
	const uint32_t k = std::min(n, capacity);
	for(uint32_t i = 0; i < k; ++i)
	{
		const ViewPortalDef& p = m->view_portals[i];
		swh_view_portal& o = out[i];
		std::memset(&o, 0, sizeof(o));
		o.kind = p.kind;
		o.blending_mode = p.blending_mode;
		o.plane[0] = p.plane.vec.x, o.plane[1] = p.plane.vec.y, o.plane[2] = p.plane.vec.z, o.plane[3] = p.plane.dist;
		for(int e = 0; e < 2; ++e)
		{
			const Plane& t = p.tex_coord_equation[e];
			o.tex_coord_equation[e][0] = t.vec.x, o.tex_coord_equation[e][1] = t.vec.y;
			o.tex_coord_equation[e][2] = t.vec.z, o.tex_coord_equation[e][3] = t.dist;
		}
		for(int v = 0; v < 4; ++v)
			o.vertices[v][0] = p.vertices[v].x, o.vertices[v][1] = p.vertices[v].y, o.vertices[v][2] = p.vertices[v].z;
		o.camera_position[0] = p.camera_position.x, o.camera_position[1] = p.camera_position.y;
		o.camera_position[2] = p.camera_position.z;
		o.camera_yaw = p.camera_yaw, o.camera_pitch = p.camera_pitch, o.camera_fov = p.camera_fov;
		for(int c = 0; c < 4; ++c)
		{
			o.transform[c * 4 + 0] = p.transform.c[c].x, o.transform[c * 4 + 1] = p.transform.c[c].y;
			o.transform[c * 4 + 2] = p.transform.c[c].z, o.transform[c * 4 + 3] = p.transform.c[c].w;
		}
		o.texture_level = p.texture_level;
		o.texture_blending_mode = p.texture_blending_mode;
		o.texture_light_scale = p.texture_light_scale;
		o.texture_light_add[0] = p.texture_light_add.x, o.texture_light_add[1] = p.texture_light_add.y;
		o.texture_light_add[2] = p.texture_light_add.z;
	}
	return int(k);
}

int swh_map_model_mesh(const swh_map* m, swh_model_mesh_data* out)
{
	if(!m || !out)
		return -1;
	const ModelMesh& mm = m->model_mesh;
	out->num_vertices = mm.num_vertices;
	out->num_triangles = uint32_t(mm.tris.size() / 3);
	out->num_frames = mm.num_frames;
	out->tc = mm.tc.data();
	out->frame_vertices = mm.abi_frame_vertices.data();
	out->triangles = mm.abi_triangles.data();
	return 0;
}

int swh_camera_path(const swh_map* m, uint32_t seed, uint32_t n, float* out)
{
	if(!m || (!out && n) || m->empty_cells.empty())
		return -1;
	Rng rng(uint64_t(seed) * 0xD1B54A32D192ED03ull + 7);
	uint32_t c = m->empty_cells[rng.below(uint32_t(m->empty_cells.size()))];
	float yaw = rng.uniform(0.0f, 6.2831853f);
	for(uint32_t i = 0; i < n; ++i)
	{
		const uint32_t x = c % m->gx, y = (c / m->gx) % m->gy, z = c / (m->gx * m->gy);
		out[i * 5 + 0] = CELL * (float(x) + 0.5f) + rng.uniform(-16.0f, 16.0f);
		out[i * 5 + 1] = CELL * (float(y) + 0.5f) + rng.uniform(-16.0f, 16.0f);
		out[i * 5 + 2] = CELL * (float(z) + 0.5f) + rng.uniform(-8.0f, 8.0f);
		out[i * 5 + 3] = yaw + rng.uniform(-0.3f, 0.3f);
		out[i * 5 + 4] = rng.uniform(-0.2f, 0.2f);
		// walk to a random empty neighbour (horizontal moves preferred), facing the move
		uint32_t options[6];
		int dirs[6];
		uint32_t no = 0;
		for(int d = 0; d < 6; ++d)
		{
			const uint32_t nc = cell_index(*m, x + uint32_t(DIRS[d][0]), y + uint32_t(DIRS[d][1]), z + uint32_t(DIRS[d][2]));
			if(m->cell[nc] == 0 && (d < 4 || rng.uniform(0.0f, 1.0f) < 0.2f))
			{
				options[no] = nc;
				dirs[no] = d;
				no++;
			}
		}
		if(no)
		{
			const uint32_t k = rng.below(no);
			c = options[k];
			if(dirs[k] < 4)
				yaw = std::atan2(float(DIRS[dirs[k]][1]), float(DIRS[dirs[k]][0]));
		}
		else
			c = m->empty_cells[rng.below(uint32_t(m->empty_cells.size()))];
	}
	return 0;
}

} // extern "C"
