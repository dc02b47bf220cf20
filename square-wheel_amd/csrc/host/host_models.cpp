// Dynamic lights, triangle models and decals of a frame.
#include <algorithm>
#include <cstring>
#include <random>

#include "host_frame_state.h"

using namespace swh;

namespace swh
{

void build_dynamic_lights(swh_frame& f, const swh_map& m, const swh_frame_config& cfg)
{
	f.lights.clear();
	f.leaf_lights.clear();
	if(cfg.num_dynamic_lights == 0 || m.empty_cells.empty())
		return;
	f.leaf_lights.assign(m.leafs.size(), std::vector<uint32_t>());
	std::mt19937_64 g(uint64_t(cfg.seed) * 0x2545F4914F6CDD1Dull + 17);
	auto uni = [&](float a, float b) { return a + (b - a) * float(double(g() >> 11) * (1.0 / 9007199254740992.0)); };
	const float CELL = 128.0f;
	for(uint32_t i = 0; i < cfg.num_dynamic_lights; ++i)
	{
		const uint32_t c = m.empty_cells[uint32_t(g() % m.empty_cells.size())];
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		const float phase = uni(0.0f, 6.2831853f);
		const V3 pos = v3(CELL * (float(x) + 0.5f) + 40.0f * std::cos(cfg.time * 1.3f + phase),
			CELL * (float(y) + 0.5f) + 40.0f * std::sin(cfg.time * 1.1f + phase), CELL * (float(z) + 0.5f) + uni(-20.0f, 20.0f));
		const float radius = uni(200.0f, 450.0f);
		sw_light l;
		l.position[0] = pos.x;
		l.position[1] = pos.y;
		l.position[2] = pos.z;
		l.radius = radius;
		l.inv_square_radius = 1.0f / (radius * radius);
		l.color[0] = uni(0.3f, 1.0f) * 15000.0f;
		l.color[1] = uni(0.3f, 1.0f) * 15000.0f;
		l.color[2] = uni(0.3f, 1.0f) * 15000.0f;
		const uint32_t index = uint32_t(f.lights.size());
		f.lights.push_back(l);
		// leaves touched by the light's sphere (the reference's dynamic_objects_index)
		const int r = int(std::ceil(radius / CELL));
		for(int dz = -r; dz <= r; ++dz)
			for(int dy = -r; dy <= r; ++dy)
				for(int dx = -r; dx <= r; ++dx)
				{
					const int cx = int(x) + dx, cy = int(y) + dy, cz = int(z) + dz;
					if(cx < 0 || cy < 0 || cz < 0 || cx >= int(m.gx) || cy >= int(m.gy) || cz >= int(m.gz))
						continue;
					const uint32_t cell = cell_index(m, uint32_t(cx), uint32_t(cy), uint32_t(cz));
					if(m.cell[cell] != 0)
						continue;
					const BBox& b = m.leaf_bbox[m.cell_leaf[cell]];
					const float qx = std::max(b.min.x, std::min(pos.x, b.max.x)) - pos.x;
					const float qy = std::max(b.min.y, std::min(pos.y, b.max.y)) - pos.y;
					const float qz = std::max(b.min.z, std::min(pos.z, b.max.z)) - pos.z;
					if(qx * qx + qy * qy + qz * qz <= radius * radius)
						f.leaf_lights[m.cell_leaf[cell]].push_back(index);
				}
	}
}

namespace
{

const float CELL = 128.0f;

struct Rng64
{
	std::mt19937_64 g;
	explicit Rng64(uint64_t seed) : g(seed) {}
	float uniform(float a, float b) { return a + (b - a) * float(double(g() >> 11) * (1.0 / 9007199254740992.0)); }
	uint32_t below(uint32_t n) { return n ? uint32_t(g() % n) : 0; }
};

// Leaves of the empty cells a world-space box overlaps (the dynamic objects index).
void leafs_of_box(const swh_map& m, V3 mn, V3 mx, std::vector<uint32_t>& out)
{
	out.clear();
	const int x0 = std::max(0, int(std::floor(mn.x / CELL))), x1 = std::min(int(m.gx) - 1, int(std::floor(mx.x / CELL)));
	const int y0 = std::max(0, int(std::floor(mn.y / CELL))), y1 = std::min(int(m.gy) - 1, int(std::floor(mx.y / CELL)));
	const int z0 = std::max(0, int(std::floor(mn.z / CELL))), z1 = std::min(int(m.gz) - 1, int(std::floor(mx.z / CELL)));
	for(int z = z0; z <= z1; ++z)
		for(int y = y0; y <= y1; ++y)
			for(int x = x0; x <= x1; ++x)
			{
				const uint32_t c = cell_index(m, uint32_t(x), uint32_t(y), uint32_t(z));
				if(m.cell[c] == 0)
					out.push_back(m.cell_leaf[c]);
			}
}

M4 rotation_z(float a)
{
	M4 r = identity();
	const float c = std::cos(a), s = std::sin(a);
	r.c[0] = v4(c, s, 0.0f, 0.0f);
	r.c[1] = v4(-s, c, 0.0f, 0.0f);
	return r;
}

bool zero_element(const LightGridElement& e)
{
	const float* p = &e.light_cube[0][0];
	for(int i = 0; i < 18; ++i)
		if(p[i] != 0.0f)
			return false;
	return e.light_direction_vector_scaled.x == 0.0f && e.light_direction_vector_scaled.y == 0.0f &&
		e.light_direction_vector_scaled.z == 0.0f && e.directional_light_color[0] == 0.0f &&
		e.directional_light_color[1] == 0.0f && e.directional_light_color[2] == 0.0f;
}

// fetch_light_from_grid (triangle_models_rendering.rs:582-700): trilinear over the 8 nearest
// samples, skipping missing (solid) ones and renormalising.
LightGridElement fetch_light_from_grid(const swh_map& m, V3 pos)
{
	LightGridElement total{};
	if(m.light_grid.empty())
		return total;
	const V3 gp = (pos - v3(CELL * 0.5f, CELL * 0.5f, CELL * 0.5f)) / CELL;
	const int gi[3] = {int(std::floor(gp.x)), int(std::floor(gp.y)), int(std::floor(gp.z))};
	const float g[3] = {gp.x, gp.y, gp.z};
	const int size[3] = {int(m.gx), int(m.gy), int(m.gz)};
	float total_factor = 0.0f;
	for(int dx = 0; dx <= 1; ++dx)
	{
		const int x = gi[0] + dx;
		if(x < 0 || x >= size[0])
			continue;
		const float fx = 1.0f - std::fabs(g[0] - float(x));
		for(int dy = 0; dy <= 1; ++dy)
		{
			const int y = gi[1] + dy;
			if(y < 0 || y >= size[1])
				continue;
			const float fy = 1.0f - std::fabs(g[1] - float(y));
			for(int dz = 0; dz <= 1; ++dz)
			{
				const int z = gi[2] + dz;
				if(z < 0 || z >= size[2])
					continue;
				const LightGridElement& s = m.light_grid[cell_index(m, uint32_t(x), uint32_t(y), uint32_t(z))];
				if(zero_element(s))
					continue;
				const float fz = 1.0f - std::fabs(g[2] - float(z));
				const float k = fx * fy * fz;
				for(int i = 0; i < 3; ++i)
				{
					for(int side = 0; side < 6; ++side)
						total.light_cube[side][i] += s.light_cube[side][i] * k;
					total.directional_light_color[i] += s.directional_light_color[i] * k;
				}
				total.light_direction_vector_scaled = total.light_direction_vector_scaled + s.light_direction_vector_scaled * k;
				total_factor += k;
			}
		}
	}
	if(total_factor <= 0.0f)
		return LightGridElement{};
	if(total_factor < 0.995f)
	{
		const float inv = 1.0f / total_factor;
		for(int i = 0; i < 3; ++i)
		{
			for(int side = 0; side < 6; ++side)
				total.light_cube[side][i] *= inv;
			total.directional_light_color[i] *= inv;
		}
		total.light_direction_vector_scaled = total.light_direction_vector_scaled * inv;
	}
	return total;
}

// get_light_cube_light (light.rs:296-340)
void light_cube_light(const float cube[6][3], V3 n, float out[3])
{
	const float comps[3] = {n.x, n.y, n.z};
	out[0] = out[1] = out[2] = 0.0f;
	for(int axis = 0; axis < 3; ++axis)
	{
		const int side = comps[axis] <= 0.0f ? axis * 2 : axis * 2 + 1;
		const float k = comps[axis] <= 0.0f ? -comps[axis] : comps[axis];
		for(int i = 0; i < 3; ++i)
			out[i] += cube[side][i] * k;
	}
}

// The two-part model light of triangle_models_rendering.rs:398-460: static grid light plus the
// dynamic lights' cube, light cubes summed, both directional components kept.
struct ModelLight
{
	float cube[6][3];
	V3 dir[2];
	float color[2][3];
};

// calculate_model_dynamic_light_cube (triangle_models_rendering.rs:494-539), shadows off.
LightGridElement model_dynamic_light(const std::vector<sw_light>& lights, const BBox& bbox, const M4& model_matrix)
{
	swh::LightCubeAccum acc;
	const V3 center = (bbox.min + bbox.max) * 0.5f;
	const V3 size = bbox.max - bbox.min;
	const float min_square_distance = dot(size, size) * 0.25f;
	V3 samples[8];
	for(int k = 0; k < 8; ++k)
	{
		const V3 corner = v3((k & 1) ? bbox.max.x : bbox.min.x, (k & 2) ? bbox.max.y : bbox.min.y, (k & 4) ? bbox.max.z : bbox.min.z);
		samples[k] = truncate(mul(model_matrix, extend((corner + center) * 0.5f, 1.0f)));
	}
	for(const sw_light& l : lights)
		for(const V3& p : samples)
		{
			const V3 to_light = v3(l.position[0], l.position[1], l.position[2]) - p;
			const float square_dist = std::max(dot(to_light, to_light), min_square_distance);
			const float inv_square_dist = 1.0f / square_dist;
			if(inv_square_dist < l.inv_square_radius)
				continue;
			const float k = inv_square_dist - l.inv_square_radius;
			const float light[3] = {l.color[0] * k, l.color[1] * k, l.color[2] * k};
			acc.add_sample(to_light, light);
		}
	acc.scale(1.0f / 8.0f);
	return acc.convert();
}

// get_vertex_light (triangle_models_rendering.rs:702-720)
void vertex_light(const ModelLight& l, V3 normal, float out[3])
{
	const V3 n = normal * (1.0f / std::sqrt(std::max(dot(normal, normal), 0.00000001f)));
	light_cube_light(l.cube, n, out);
	for(int c = 0; c < 2; ++c)
	{
		const float d = std::max(dot(n, l.dir[c]), 0.0f);
		for(int i = 0; i < 3; ++i)
			out[i] += l.color[c][i] * d;
	}
}

int32_t f32_to_fixed16(float x) { return saturating_f32_to_i32(x * 65536.0f); }

void emit_triangle(swh_frame& f, const swh_map& m, const float pts[3][7], uint32_t level, uint32_t blend, const int32_t clip_rect[4])
{
	const sw_texture_level& tl = m.levels[level];
	sw_draw d;
	std::memset(&d, 0, sizeof(d));
	d.kind = SW_DRAW_TRIANGLE;
	d.blend = blend;
	d.first_vertex_int = uint32_t(f.vdata.size());
	d.num_vertices = 3;
	d.texture_offset = tl.offset;
	d.texture_size[0] = int32_t(tl.width);
	d.texture_size[1] = int32_t(tl.height);
	for(int i = 0; i < 4; ++i)
		d.clip_rect[i] = clip_rect[i];
	for(int v = 0; v < 3; ++v)
		for(int k = 0; k < 7; ++k)
			f.vdata.push_back(f32_to_fixed16(pts[v][k]));
	f.draws.push_back(d);
	f.num_triangles++;
}

struct ModelVertex2d
{
	V2 pos;
	float tc[2];
	float light[3];
};

template <class VT>
VT lerp_vertex(const VT& a, const VT& b, float k0, float k1)
{
	// v0 * k1 - v1 * k0 with k0 = dist0 / (dist1 - dist0), k1 = dist1 / (dist1 - dist0)
	VT r;
	r.pos = a.pos * k1 - b.pos * k0;
	for(int i = 0; i < 2; ++i)
		r.tc[i] = a.tc[i] * k1 - b.tc[i] * k0;
	for(int i = 0; i < 3; ++i)
		r.light[i] = a.light[i] * k1 - b.light[i] * k0;
	return r;
}

// clip_3d_model_polygon_by_plane / clip_2d_model_polygon (clipping.rs:203-330): keeps the part
// with dist(v) >= 0.
template <class VT, class DistF>
uint32_t clip_model_polygon(const VT* poly, uint32_t n, DistF dist_of, VT* out, uint32_t cap)
{
	if(n == 0)
		return 0;
	const VT* prev = &poly[n - 1];
	float prev_dist = dist_of(*prev);
	uint32_t count = 0;
	for(uint32_t i = 0; i < n && count < cap; ++i)
	{
		const VT& v = poly[i];
		const float dist = dist_of(v);
		if(dist > 0.0f)
		{
			if(prev_dist < 0.0f)
			{
				const float sum = dist - prev_dist;
				out[count++] = lerp_vertex(*prev, v, prev_dist / sum, dist / sum);
				if(count == cap)
					break;
			}
			out[count++] = v;
		}
		else if(dist == 0.0f)
			out[count++] = v;
		else if(prev_dist > 0.0f)
		{
			const float sum = dist - prev_dist;
			out[count++] = lerp_vertex(*prev, v, prev_dist / sum, dist / sum);
		}
		prev = &v;
		prev_dist = dist;
	}
	return count;
}

// draw_mesh (partial_renderer.rs:3525-3718)
void draw_mesh(swh_frame& f, const swh_map& m, const ClippingPolygon& bounds, const Plane* leaf_planes, uint32_t num_leaf_planes,
	const VisibleModel& vm, const int32_t clip_rect[4])
{
	ClippingPolygon mesh_clip{{vm.clip[0][0], vm.clip[0][1]}, {vm.clip[1][0], vm.clip[1][1]}, {vm.clip[2][0], vm.clip[2][1]},
		{vm.clip[3][0], vm.clip[3][1]}};
	{
		ClippingPolygon c = mesh_clip;
		c.intersect(bounds);
		if(!c.valid())
			return;
	}
	const ModelEntity& e = f.models.entities[vm.entity];
	Plane planes3d[1 + 6];
	uint32_t n3 = 0;
	Plane candidates[1 + 6];
	candidates[0] = Plane{v3(0.0f, 0.0f, 1.0f), Z_NEAR};
	for(uint32_t i = 0; i < num_leaf_planes; ++i)
		candidates[1 + i] = leaf_planes[i];
	for(uint32_t i = 0; i < 1 + num_leaf_planes; ++i)
	{
		uint32_t front = 0;
		for(const V3& v : vm.bbox_vertices)
			if(dot(candidates[i].vec, v) >= candidates[i].dist)
				++front;
		if(front == 8)
			continue;
		if(front == 0)
			return;
		planes3d[n3++] = candidates[i];
	}
	// 2D box planes of the leaf bounds that cut into the model's screen box
	const V3 mesh_box[4] = {v3(1.0f, 0.0f, mesh_clip.x.min), v3(-1.0f, 0.0f, -mesh_clip.x.max), v3(0.0f, 1.0f, mesh_clip.y.min),
		v3(0.0f, -1.0f, -mesh_clip.y.max)};
	const V3 cur_box[4] = {v3(1.0f, 0.0f, bounds.x.min), v3(-1.0f, 0.0f, -bounds.x.max), v3(0.0f, 1.0f, bounds.y.min),
		v3(0.0f, -1.0f, -bounds.y.max)};
	V3 planes2d[4];
	uint32_t n2 = 0;
	for(int i = 0; i < 4; ++i)
		if(cur_box[i].z > mesh_box[i].z)
			planes2d[n2++] = cur_box[i];

	// the run: the mesh's per-leaf decisions, at its place in the draw list
	ModelFrameData& md = f.models;
	{
		sw_mesh_run run;
		std::memset(&run, 0, sizeof(run));
		run.count_index = uint32_t(&vm - md.visible.data());
		run.first_vertex = vm.first_vertex;
		run.num_vertices = vm.num_vertices;
		run.first_triangle = vm.first_triangle;
		run.max_triangles = vm.num_triangles;
		run.num_planes3d = n3;
		run.num_planes2d = n2;
		run.blend = e.blend;
		for(uint32_t p = 0; p < n3; ++p)
		{
			run.planes3d[p][0] = planes3d[p].vec.x, run.planes3d[p][1] = planes3d[p].vec.y;
			run.planes3d[p][2] = planes3d[p].vec.z, run.planes3d[p][3] = planes3d[p].dist;
		}
		for(uint32_t p = 0; p < n2; ++p)
			run.planes2d[p][0] = planes2d[p].x, run.planes2d[p][1] = planes2d[p].y, run.planes2d[p][2] = planes2d[p].z;
		const sw_texture_level& tl = m.levels[vm.level];
		run.texture_offset = tl.offset;
		run.texture_size[0] = int32_t(tl.width);
		run.texture_size[1] = int32_t(tl.height);
		for(int i = 0; i < 4; ++i)
			run.clip_rect[i] = clip_rect[i];
		md.runs.push_back(run);
		md.run_positions.push_back(uint32_t(f.draws.size()));
		md.run_num_draws.push_back(0);
	}
	if(md.device)
		return;
	const size_t draws_before = f.draws.size();

	const ModelVertex3d* verts = &f.models.vertices[vm.first_vertex];
	const uint32_t* tris = &f.models.triangles[vm.first_triangle * 3];
	for(uint32_t t = 0; t < vm.num_triangles; ++t)
	{
		ModelVertex3d c0[MAX_VERTICES], c1[MAX_VERTICES];
		for(int k = 0; k < 3; ++k)
			c0[k] = verts[tris[t * 3 + k]];
		uint32_t n = 3;
		ModelVertex3d* src = c0;
		ModelVertex3d* dst = c1;
		bool culled = false;
		for(uint32_t p = 0; p < n3 && !culled; ++p)
		{
			const Plane pl = planes3d[p];
			n = clip_model_polygon(src, n, [&](const ModelVertex3d& v) { return dot(pl.vec, v.pos) - pl.dist; }, dst, MAX_VERTICES);
			culled = n < 3;
			std::swap(src, dst);
		}
		if(culled)
			continue;
		ModelVertex2d p0[MAX_VERTICES], p1[MAX_VERTICES];
		for(uint32_t k = 0; k < n; ++k)
		{
			p0[k].pos = v2(src[k].pos.x, src[k].pos.y) / src[k].pos.z;
			p0[k].tc[0] = src[k].tc[0];
			p0[k].tc[1] = src[k].tc[1];
			for(int i = 0; i < 3; ++i)
				p0[k].light[i] = src[k].light[i];
		}
		ModelVertex2d* s2 = p0;
		ModelVertex2d* d2 = p1;
		for(uint32_t p = 0; p < n2 && !culled; ++p)
		{
			const V3 line = planes2d[p];
			n = clip_model_polygon(
				s2, n, [&](const ModelVertex2d& v) { return v.pos.x * line.x + v.pos.y * line.y - line.z; }, d2, MAX_VERTICES);
			culled = n < 3;
			std::swap(s2, d2);
		}
		if(culled)
			continue;
		for(uint32_t k = 0; k + 2 < n; ++k)
		{
			const ModelVertex2d* tri[3] = {&s2[0], &s2[k + 1], &s2[k + 2]};
			float pts[3][7];
			for(int v = 0; v < 3; ++v)
			{
				pts[v][0] = tri[v]->pos.x;
				pts[v][1] = tri[v]->pos.y;
				pts[v][2] = tri[v]->tc[0];
				pts[v][3] = tri[v]->tc[1];
				pts[v][4] = tri[v]->light[0];
				pts[v][5] = tri[v]->light[1];
				pts[v][6] = tri[v]->light[2];
			}
			emit_triangle(f, m, pts, vm.level, e.blend, clip_rect);
		}
	}
	md.run_num_draws.back() = uint32_t(f.draws.size() - draws_before);
}

// get_polygon_lightmap_light (partial_renderer.rs:3807-3875): bilinear fetch of the polygon's
// simple lightmap.
void polygon_lightmap_light(const swh_map& m, const Polygon& polygon, V2 coord, float out[3])
{
	out[0] = out[1] = out[2] = 0.0f;
	if(polygon.lightmap_data_offset == 0 || m.lm_simple.empty())
		return;
	const int32_t lw = ((polygon.tex_coord_max[0] + int32_t(LIGHTMAP_SCALE) - 1) >> LIGHTMAP_SCALE_LOG2) -
		(polygon.tex_coord_min[0] >> LIGHTMAP_SCALE_LOG2) + 1;
	const int32_t lh = ((polygon.tex_coord_max[1] + int32_t(LIGHTMAP_SCALE) - 1) >> LIGHTMAP_SCALE_LOG2) -
		(polygon.tex_coord_min[1] >> LIGHTMAP_SCALE_LOG2) + 1;
	const int32_t ci[2] = {saturating_f32_to_i32(std::floor(coord.x)), saturating_f32_to_i32(std::floor(coord.y))};
	float total_factor = 0.0f;
	for(int dy = 0; dy <= 1; ++dy)
	{
		const int32_t y = ci[1] + dy;
		if(y < 0 || y >= lh)
			continue;
		const float fy = 1.0f - std::fabs(coord.y - float(y));
		for(int dx = 0; dx <= 1; ++dx)
		{
			const int32_t x = ci[0] + dx;
			if(x < 0 || x >= lw)
				continue;
			const float fx = 1.0f - std::fabs(coord.x - float(x));
			const float* l = &m.lm_simple[(uint64_t(polygon.lightmap_data_offset) + uint64_t(x + y * lw)) * 3];
			const float k = fx * fy;
			for(int i = 0; i < 3; ++i)
				out[i] += k * l[i];
			total_factor += k;
		}
	}
	if(total_factor < 1.0f && total_factor > 0.0f)
	{
		const float inv = 1.0f / total_factor;
		for(int i = 0; i < 3; ++i)
			out[i] *= inv;
	}
}

struct DecalDraw
{
	const DecalEntity* decal;
	const Polygon* polygon;
	DepthEq de;
	TcEq te;
	TcEq lm_eq;
	float lm_shift[2];
	uint32_t level;
	float tex_size[2];
	float dynamic_light[3];
};

// subdivide_and_draw_decal_triangle (partial_renderer.rs:2940-3096)
void decal_triangle(swh_frame& f, const swh_map& m, const DecalDraw& dd, const V2 points[3], uint32_t depth, const int32_t clip_rect[4])
{
	V2 min_p = points[0], max_p = points[0];
	float min_inv_z = dd.de.sample(points[0]), max_inv_z = min_inv_z;
	for(int i = 1; i < 3; ++i)
	{
		const float iz = dd.de.sample(points[i]);
		if(iz < min_inv_z)
		{
			min_inv_z = iz;
			min_p = points[i];
		}
		if(iz > max_inv_z)
		{
			max_inv_z = iz;
			max_p = points[i];
		}
	}
	if(depth >= 2 || affine_may_be_used(dd.de, dd.te, min_p, max_p))
	{
		const float eps = 1.0f / 32.0f;
		float pts[3][7];
		for(int v = 0; v < 3; ++v)
		{
			const V2 p = points[v];
			const float z = 1.0f / dd.de.sample(p);
			float light[3] = {dd.decal->light_add[0], dd.decal->light_add[1], dd.decal->light_add[2]};
			if(dd.decal->lightmap_light_scale > 0.0f)
			{
				const V2 lc = v2(z * (dd.lm_eq.dx[0] * p.x + dd.lm_eq.dy[0] * p.y + dd.lm_eq.k[0]) + dd.lm_shift[0],
					z * (dd.lm_eq.dx[1] * p.x + dd.lm_eq.dy[1] * p.y + dd.lm_eq.k[1]) + dd.lm_shift[1]);
				float lm[3];
				polygon_lightmap_light(m, *dd.polygon, lc, lm);
				for(int i = 0; i < 3; ++i)
				{
					light[i] += lm[i] * dd.decal->lightmap_light_scale;
					light[i] += dd.dynamic_light[i];
				}
			}
			pts[v][0] = p.x;
			pts[v][1] = p.y;
			for(int i = 0; i < 2; ++i)
			{
				const float tc = z * (dd.te.dx[i] * p.x + dd.te.dy[i] * p.y + dd.te.k[i]);
				pts[v][2 + i] = std::min(std::max(tc, eps), dd.tex_size[i] - eps);
			}
			pts[v][4] = light[0];
			pts[v][5] = light[1];
			pts[v][6] = light[2];
		}
		emit_triangle(f, m, pts, dd.level, dd.decal->blend, clip_rect);
		return;
	}
	const V2 c0 = (points[0] + points[1]) / 2.0f, c1 = (points[1] + points[2]) / 2.0f, c2 = (points[2] + points[0]) / 2.0f;
	const V2 sub[4][3] = {{points[0], c0, c2}, {points[1], c1, c0}, {points[2], c2, c1}, {c0, c1, c2}};
	for(const auto& t : sub)
		decal_triangle(f, m, dd, t, depth + 1, clip_rect);
}

// draw_polygon_decals (partial_renderer.rs:2794-2938)
void draw_polygon_decals(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const V3 clip_planes[8], uint32_t pi,
	const std::vector<uint32_t>& decals, const int32_t clip_rect[4])
{
	const swh_frame::PolyData& pd = f.poly[pi];
	if(pd.visible_frame != f.frame_number)
		return;
	const Polygon& polygon = m.polygons[pi];
	if(m.materials[polygon.texture].sky)
		return;
	static const float CUBE_SIDES[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
	const V3 normal = f.basis[pi].normal;
	for(uint32_t di : decals)
	{
		const DecalEntity& decal = f.models.decals[di];
		const ModelFrameData::DecalInfo& info = f.models.decal_info[di];
		// both vectors are unit: the dot is the cosine; skip back faces and steep slopes
		if(dot(normal, decal.axis[0] * -1.0f) < 0.1f)
			continue;
		V3 a[MAX_VERTICES], b[MAX_VERTICES];
		uint32_t n = std::min(polygon.num_vertices, MAX_VERTICES);
		for(uint32_t k = 0; k < n; ++k)
			a[k] = f.vertices_transformed[polygon.first_vertex + k];
		V3* src = a;
		V3* dst = b;
		bool culled = false;
		for(int side = 0; side < 6 && !culled; ++side)
		{
			const V4 pv = mul(info.camera_planes_matrix, v4(CUBE_SIDES[side][0], CUBE_SIDES[side][1], CUBE_SIDES[side][2], 1.0f));
			const V3 vec = truncate(pv);
			const float dist = -pv.w;
			// clip_3d_polygon_by_plane (clipping.rs:53-103)
			const V3* prev = &src[n - 1];
			float prev_d = dot(vec, *prev);
			uint32_t count = 0;
			for(uint32_t k = 0; k < n && count < MAX_VERTICES; ++k)
			{
				const V3& v = src[k];
				const float d = dot(vec, v);
				auto cut = [&]() {
					const float d0 = prev_d - dist, d1 = d - dist, sum = d1 - d0;
					return *prev * (d1 / sum) - v * (d0 / sum);
				};
				if(d > dist)
				{
					if(prev_d < dist)
					{
						dst[count++] = cut();
						if(count == MAX_VERTICES)
							break;
					}
					dst[count++] = v;
				}
				else if(d == dist)
					dst[count++] = v;
				else if(prev_d > dist)
					dst[count++] = cut();
				prev = &v;
				prev_d = d;
			}
			n = count;
			culled = n < 3;
			std::swap(src, dst);
		}
		if(culled)
			continue;
		V2 pts[MAX_VERTICES];
		n = project_and_clip_polygon(clip_planes, src, n, pts);
		if(n < 3)
			continue;
		const sw_texture_level& l0 = m.levels[decal.level];
		const V4 basis[2] = {v4(0.0f, -0.5f * float(l0.width), 0.0f, 0.5f * float(l0.width)),
			v4(0.0f, 0.0f, -0.5f * float(l0.height), 0.5f * float(l0.height))};
		const V4 tcb[2] = {mul(info.camera_planes_matrix, basis[0]), mul(info.camera_planes_matrix, basis[1])};
		DecalDraw dd;
		dd.decal = &decal;
		dd.polygon = &polygon;
		dd.de = pd.de;
		const TcEq te = tc_from_depth_and_planes(pd.de, tcb);
		const uint32_t mip = calculate_mip(pts, n, pd.de, te, cfg.textures_mip_bias);
		dd.level = decal.level + mip;
		dd.tex_size[0] = float(m.levels[dd.level].width);
		dd.tex_size[1] = float(m.levels[dd.level].height);
		dd.te = scale(te, 1.0f / float(1u << mip));
		const float lm_scale = float(1u << pd.mip) / float(LIGHTMAP_SCALE);
		for(int i = 0; i < 2; ++i)
			dd.lm_shift[i] = float(pd.tc_min[i]) * lm_scale - float(polygon.tex_coord_min[i] >> LIGHTMAP_SCALE_LOG2);
		dd.lm_eq = scale(pd.te, lm_scale);
		light_cube_light(info.dynamic_light.light_cube, normal, dd.dynamic_light);
		const float dl = std::max(dot(normal, info.dynamic_light.light_direction_vector_scaled), 0.0f);
		for(int i = 0; i < 3; ++i)
			dd.dynamic_light[i] += dl * info.dynamic_light.directional_light_color[i];
		for(uint32_t t = 0; t + 2 < n; ++t)
		{
			const V2 tri[3] = {pts[0], pts[t + 1], pts[t + 2]};
			decal_triangle(f, m, dd, tri, 0, clip_rect);
		}
	}
}

} // namespace

void grid_surface_light(const swh_map& m, V3 pos, V3 n, float out[3])
{
	const LightGridElement g = fetch_light_from_grid(m, pos);
	light_cube_light(g.light_cube, n, out);
	const float d = std::max(dot(n, g.light_direction_vector_scaled), 0.0f);
	for(int i = 0; i < 3; ++i)
		out[i] += g.directional_light_color[i] * d;
}

// Model entities of the frame: seeded placement (stable along a camera path), animation and spin
// from the frame time.
void prepare_models(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const M4& view, const M4& planes)
{
	ModelFrameData& md = f.models;
	md.entities.clear();
	md.visible.clear();
	md.vertices.clear();
	md.triangles.clear();
	md.leaf_models.clear();
	md.view_models.clear();
	md.device = cfg.device_models != 0;
	md.instances.clear();
	md.runs.clear();
	md.run_positions.clear();
	md.run_num_draws.clear();
	md.prepared_vertices.clear();
	md.prepared_triangles.clear();
	md.prepared_counts.clear();
	md.vertex_capacity = md.triangle_capacity = 0;
	const ModelMesh& mesh = m.model_mesh;
	if((cfg.num_models == 0 && cfg.view_model == 0) || m.model_levels.empty() || m.empty_cells.empty() || mesh.num_vertices == 0)
		return;
	md.leaf_models.assign(m.leafs.size(), std::vector<uint32_t>());
	Rng64 rng(uint64_t(cfg.seed) * 0x9E3779B97F4A7C15ull + 101);
	for(uint32_t i = 0; i < cfg.num_models; ++i)
	{
		ModelEntity e;
		const uint32_t c = m.empty_cells[rng.below(uint32_t(m.empty_cells.size()))];
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		e.pos = v3(CELL * (float(x) + rng.uniform(0.4f, 0.6f)), CELL * (float(y) + rng.uniform(0.4f, 0.6f)),
			CELL * (float(z) + rng.uniform(0.35f, 0.5f)));
		const float spin = rng.uniform(-1.0f, 1.0f), phase = rng.uniform(0.0f, 8.0f);
		e.yaw = rng.uniform(0.0f, 6.2831853f) + spin * cfg.time;
		const float t = cfg.time * 10.0f + phase;
		const float ft = std::floor(t);
		e.frames[0] = uint32_t(int64_t(ft) % int64_t(mesh.num_frames));
		e.frames[1] = (e.frames[0] + 1) % mesh.num_frames;
		e.lerp = 1.0f - (t - ft);
		e.skin = m.model_levels[i % m.model_levels.size()];
		e.blend = (i % 8 == 7) ? SW_BLEND_ADDITIVE : SW_BLEND_NONE;
		md.entities.push_back(e);
	}
	M4 inv_view;
	if(cfg.view_model && invert(view, inv_view))
	{
		// held at the lower right of the screen, 48 units in front of the camera
		const float z = 48.0f;
		ModelEntity e;
		e.pos = truncate(mul(inv_view, v4(0.72f * float(cfg.width) * z, 0.8f * float(cfg.height) * z, z, 1.0f)));
		e.yaw = 0.6f * cfg.time;
		const float t = cfg.time * 10.0f;
		const float ft = std::floor(t);
		e.frames[0] = uint32_t(int64_t(ft) % int64_t(mesh.num_frames));
		e.frames[1] = (e.frames[0] + 1) % mesh.num_frames;
		e.lerp = 1.0f - (t - ft);
		e.skin = m.model_levels[0];
		e.blend = SW_BLEND_NONE;
		e.view_model = true;
		md.entities.push_back(e);
	}

	std::vector<uint32_t> leafs;
	for(uint32_t ei = 0; ei < md.entities.size(); ++ei)
	{
		const ModelEntity& e = md.entities[ei];
		const M4 model_matrix = mul(from_translation(e.pos), rotation_z(e.yaw));
		const M4 mv = mul(view, model_matrix);
		BBox bbox = mesh.frame_bbox[e.frames[0]];
		const BBox& b1 = mesh.frame_bbox[e.frames[1]];
		bbox.min = v3(std::min(bbox.min.x, b1.min.x), std::min(bbox.min.y, b1.min.y), std::min(bbox.min.z, b1.min.z));
		bbox.max = v3(std::max(bbox.max.x, b1.max.x), std::max(bbox.max.y, b1.max.y), std::max(bbox.max.z, b1.max.z));
		VisibleModel vm;
		V3 world_min = v3(1e30f, 1e30f, 1e30f), world_max = v3(-1e30f, -1e30f, -1e30f);
		for(int k = 0; k < 8; ++k)
		{
			const V3 corner = v3((k & 1) ? bbox.max.x : bbox.min.x, (k & 2) ? bbox.max.y : bbox.min.y, (k & 4) ? bbox.max.z : bbox.min.z);
			vm.bbox_vertices[k] = view_transform(mv, corner);
			const V3 w = truncate(mul(model_matrix, extend(corner, 1.0f)));
			world_min = v3(std::min(world_min.x, w.x), std::min(world_min.y, w.y), std::min(world_min.z, w.z));
			world_max = v3(std::max(world_max.x, w.x), std::max(world_max.y, w.y), std::max(world_max.z, w.z));
		}
		// calculate_triangle_model_screen_polygon (triangle_models_rendering.rs:214-272)
		bool have = false;
		uint32_t front = 0;
		ClippingPolygon sp{};
		for(const V3& v : vm.bbox_vertices)
			if(v.z > 0.0f)
			{
				++front;
				const V2 p = v2(v.x, v.y) / v.z;
				if(have)
					sp.extend_point(p);
				else
					sp = ClippingPolygon::from_point(p);
				have = true;
			}
		if(!have)
			continue;
		if(front != 8)
		{
			const float ZN = 1.0f / 4096.0f;
			for(int a = 0; a < 7; ++a)
				for(int b = a + 1; b < 8; ++b)
				{
					const V3& va = vm.bbox_vertices[a];
					const V3& vb = vm.bbox_vertices[b];
					if((va.z <= ZN && vb.z <= ZN) || (va.z >= ZN && vb.z >= ZN))
						continue;
					const V3 v = line_z_intersection(va, vb, ZN);
					sp.extend_point(v2(v.x, v.y) / v.z);
				}
		}
		// visible if one of its leaves is visible inside the model's screen box (view models always)
		leafs.clear();
		if(!e.view_model)
			leafs_of_box(m, world_min, world_max, leafs);
		bool visible = e.view_model;
		for(uint32_t l : leafs)
		{
			if(f.leaf_visible_frame[l] != f.frame_number)
				continue;
			ClippingPolygon lb = f.leaf_bounds[l];
			lb.intersect(sp);
			if(lb.valid())
			{
				visible = true;
				break;
			}
		}
		if(!visible)
			continue;
		vm.entity = ei;
		vm.clip[0][0] = sp.x.min, vm.clip[0][1] = sp.x.max;
		vm.clip[1][0] = sp.y.min, vm.clip[1][1] = sp.y.max;
		vm.clip[2][0] = sp.xpy.min, vm.clip[2][1] = sp.xpy.max;
		vm.clip[3][0] = sp.xmy.min, vm.clip[3][1] = sp.xmy.max;
		const V3 center_cam = view_transform(mv, (bbox.min + bbox.max) * 0.5f);
		vm.depth = center_cam.z;
		vm.bbox = bbox;
		vm.view = mv;
		M4 world_planes;
		if(!invert(transpose(model_matrix), world_planes))
			world_planes = identity();
		vm.planes = mul(planes, world_planes);

		// calculate_triangle_model_texture_mip (triangle_models_rendering.rs:274-322)
		uint32_t mip = 0;
		{
			const sw_texture_level& l0 = m.levels[e.skin];
			const V3 c = (bbox.min + bbox.max) * 0.5f, hs = (bbox.max - bbox.min) * 0.5f;
			const float cube = (hs.x + hs.y + hs.z) / 3.0f;
			float min_x = 0, max_x = 0, min_y = 0, max_y = 0;
			bool behind = false;
			for(int k = 0; k < 8; ++k)
			{
				const V3 corner = v3((k & 1) ? c.x + cube : c.x - cube, (k & 2) ? c.y + cube : c.y - cube, (k & 4) ? c.z + cube : c.z - cube);
				const V3 pr = view_transform(mv, corner);
				if(pr.z <= 0.0f)
				{
					behind = true;
					break;
				}
				const V2 q = v2(pr.x, pr.y) / pr.z;
				if(k == 0)
				{
					min_x = max_x = q.x;
					min_y = max_y = q.y;
				}
				min_x = std::min(min_x, q.x);
				max_x = std::max(max_x, q.x);
				min_y = std::min(min_y, q.y);
				max_y = std::max(max_y, q.y);
			}
			if(!behind)
			{
				const float max_dim = std::max(max_x - min_x, max_y - min_y);
				if(max_dim <= 0.0f)
					mip = MAX_MIP;
				else
				{
					const float tpp = float(std::max(l0.width, l0.height)) / max_dim;
					const float mf = std::floor(std::log2(tpp) + cfg.textures_mip_bias);
					mip = uint32_t(std::min(float(MAX_MIP), std::max(0.0f, mf)));
				}
			}
		}
		vm.level = e.skin + mip;
		const float tsize[2] = {float(m.levels[vm.level].width), float(m.levels[vm.level].height)};

		// lights: static (grid) + dynamic cube (ModelLighting::Default)
		ModelLight ml;
		const LightGridElement st = fetch_light_from_grid(m, e.pos);
		const LightGridElement dy = model_dynamic_light(f.lights, bbox, model_matrix);
		for(int side = 0; side < 6; ++side)
			for(int i = 0; i < 3; ++i)
				ml.cube[side][i] = st.light_cube[side][i] + dy.light_cube[side][i];
		ml.dir[0] = st.light_direction_vector_scaled;
		ml.dir[1] = dy.light_direction_vector_scaled;
		for(int i = 0; i < 3; ++i)
		{
			ml.color[0][i] = st.directional_light_color[i];
			ml.color[1][i] = dy.directional_light_color[i];
		}

		// the model as an instance of the device preparation (sw_prepare_models' arguments:
		// animation point, matrices, tc scale, ModelLightData; outputs in capacity order)
		{
			sw_model_instance in;
			std::memset(&in, 0, sizeof(in));
			in.mesh = 0;
			in.frames[0] = e.frames[0];
			in.frames[1] = e.frames[1];
			in.lerp = e.lerp;
			for(int c = 0; c < 4; ++c)
				for(int r = 0; r < 4; ++r)
				{
					in.model_matrix[c * 4 + r] = model_matrix.at(r, c);
					in.model_view_matrix[c * 4 + r] = mv.at(r, c);
				}
			in.tc_scale[0] = tsize[0];
			in.tc_scale[1] = tsize[1];
			for(int side = 0; side < 6; ++side)
				for(int i = 0; i < 3; ++i)
					in.light_cube[side][i] = ml.cube[side][i];
			for(int d = 0; d < 2; ++d)
			{
				in.light_vector_scaled[d][0] = ml.dir[d].x;
				in.light_vector_scaled[d][1] = ml.dir[d].y;
				in.light_vector_scaled[d][2] = ml.dir[d].z;
				for(int i = 0; i < 3; ++i)
					in.light_color[d][i] = ml.color[d][i];
			}
			in.first_vertex = uint32_t(md.vertex_capacity);
			in.first_triangle = uint32_t(md.triangle_capacity);
			md.vertex_capacity += mesh.num_vertices;
			md.triangle_capacity += mesh.tris.size() / 3;
			md.instances.push_back(in);
			if(md.device)
			{
				// the device animates, lights, culls and sorts: the runs draw up to every triangle
				vm.first_vertex = in.first_vertex;
				vm.num_vertices = mesh.num_vertices;
				vm.first_triangle = in.first_triangle;
				vm.num_triangles = uint32_t(mesh.tris.size() / 3);
				const uint32_t vis = uint32_t(md.visible.size());
				md.visible.push_back(vm);
				if(e.view_model)
					md.view_models.push_back(vis);
				for(uint32_t l : leafs)
					md.leaf_models[l].push_back(vis);
				continue;
			}
		}

		// animate_and_transform_triangle_mesh_vertices (triangle_models_rendering.rs:7-200):
		// the normals matrix of a rotation is the rotation itself
		vm.first_vertex = uint32_t(md.vertices.size());
		vm.num_vertices = mesh.num_vertices;
		const float l0 = std::max(0.0f, std::min(1.0f, e.lerp)), l1 = 1.0f - l0;
		const bool lerp = l0 > 0.01f && l0 < 0.99f;
		const V3* p0 = &mesh.pos[size_t(e.frames[0]) * mesh.num_vertices];
		const V3* p1 = &mesh.pos[size_t(e.frames[1]) * mesh.num_vertices];
		const V3* n0 = &mesh.normal[size_t(e.frames[0]) * mesh.num_vertices];
		const V3* n1 = &mesh.normal[size_t(e.frames[1]) * mesh.num_vertices];
		for(uint32_t v = 0; v < mesh.num_vertices; ++v)
		{
			V3 pos, nrm;
			if(lerp)
			{
				pos = p0[v] * l0 + p1[v] * l1;
				nrm = n0[v] * l0 + n1[v] * l1;
			}
			else
			{
				pos = l0 > l1 ? p0[v] : p1[v];
				nrm = l0 > l1 ? n0[v] : n1[v];
			}
			const V3 nt = truncate(mul(model_matrix, extend(nrm, 0.0f)));
			ModelVertex3d mv3;
			mv3.pos = view_transform(mv, pos);
			mv3.tc[0] = mesh.tc[v * 2] * tsize[0];
			mv3.tc[1] = mesh.tc[v * 2 + 1] * tsize[1];
			vertex_light(ml, nt, mv3.light);
			md.vertices.push_back(mv3);
		}
		// reject_triangle_model_back_faces + sort_model_triangles (:324-391)
		const ModelVertex3d* tv = &md.vertices[vm.first_vertex];
		vm.first_triangle = uint32_t(md.triangles.size() / 3);
		std::vector<std::pair<float, uint32_t>> order;
		for(size_t k = 0; k < mesh.tris.size(); k += 3)
		{
			// the mesh is wound outward counter-clockwise; the camera sees its front faces clockwise
			const V3 a = tv[mesh.tris[k]].pos, b = tv[mesh.tris[k + 2]].pos, c = tv[mesh.tris[k + 1]].pos;
			const V3 vec = cross(b - a, c - b);
			if(dot(vec, a) > 0.0f)
				order.push_back({std::max(a.z, std::max(b.z, c.z)), uint32_t(k)});
		}
		std::stable_sort(order.begin(), order.end(), [](const std::pair<float, uint32_t>& x, const std::pair<float, uint32_t>& y) {
			return x.first > y.first;
		});
		for(const auto& o : order)
		{
			md.triangles.push_back(mesh.tris[o.second]);
			md.triangles.push_back(mesh.tris[o.second + 2]);
			md.triangles.push_back(mesh.tris[o.second + 1]);
		}
		vm.num_triangles = uint32_t(order.size());
		const uint32_t vis = uint32_t(md.visible.size());
		md.visible.push_back(vm);
		if(e.view_model)
			md.view_models.push_back(vis);
		for(uint32_t l : leafs)
			md.leaf_models[l].push_back(vis);
	}
	if(!md.device)
	{
		// the host-prepared models in the device preparation's output layout
		md.prepared_vertices.resize(md.vertices.size());
		for(size_t i = 0; i < md.vertices.size(); ++i)
		{
			const ModelVertex3d& v = md.vertices[i];
			sw_model_vertex& o = md.prepared_vertices[i];
			o.pos[0] = v.pos.x, o.pos[1] = v.pos.y, o.pos[2] = v.pos.z;
			o.tc[0] = v.tc[0], o.tc[1] = v.tc[1];
			o.light[0] = v.light[0], o.light[1] = v.light[1], o.light[2] = v.light[2];
		}
		md.prepared_triangles.assign(md.triangles.begin(), md.triangles.end());
		for(const VisibleModel& vm : md.visible)
			md.prepared_counts.push_back(vm.num_triangles);
	}
}

// Decals of the frame (seeded, stable along a camera path) and their camera-space planes and
// dynamic light (partial_renderer.rs:665-736).
void prepare_decals(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const M4& planes)
{
	ModelFrameData& md = f.models;
	md.decals.clear();
	md.decal_info.clear();
	md.leaf_decals.clear();
	if(cfg.num_decals == 0 || m.decal_levels.empty() || m.polygons.empty())
		return;
	md.leaf_decals.assign(m.leafs.size(), std::vector<uint32_t>());
	Rng64 rng(uint64_t(cfg.seed) * 0xD1B54A32D192ED03ull + 7);
	std::vector<uint32_t> leafs;
	for(uint32_t i = 0; i < cfg.num_decals; ++i)
	{
		uint32_t pi = 0;
		for(int tries = 0; tries < 64; ++tries)
		{
			pi = rng.below(uint32_t(m.polygons.size()));
			if(!m.materials[m.polygons[pi].texture].sky)
				break;
		}
		const Polygon& p = m.polygons[pi];
		if(m.materials[p.texture].sky)
			continue;
		const V3 n = p.plane.vec / length(p.plane.vec);
		V3 c = v3(0, 0, 0);
		for(uint32_t k = 0; k < p.num_vertices; ++k)
			c = c + m.vertices[p.first_vertex + k];
		c = c / float(p.num_vertices);
		DecalEntity d;
		// decal x points into the wall, y and z span the wall with a random roll
		d.axis[0] = n * -1.0f;
		const V3 helper = std::fabs(n.z) < 0.9f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
		V3 t1 = cross(helper, d.axis[0]);
		t1 = t1 / length(t1);
		const V3 t2 = cross(d.axis[0], t1);
		const float roll = rng.uniform(0.0f, 6.2831853f);
		d.axis[1] = t1 * std::cos(roll) + t2 * std::sin(roll);
		d.axis[2] = cross(d.axis[0], d.axis[1]);
		const float off_a = rng.uniform(-40.0f, 40.0f), off_b = rng.uniform(-40.0f, 40.0f);
		d.pos = c + t1 * off_a + t2 * off_b + n * 2.0f;
		const float s = rng.uniform(10.0f, 26.0f);
		d.scale = v3(8.0f, s, s * rng.uniform(0.7f, 1.3f));
		d.level = m.decal_levels[i % m.decal_levels.size()];
		d.blend = SW_BLEND_ALPHA_BLEND;
		d.lightmap_light_scale = 1.0f;
		d.light_add[0] = d.light_add[1] = d.light_add[2] = 0.0f;
		const uint32_t index = uint32_t(md.decals.size());
		md.decals.push_back(d);

		// decal matrix: translation * rotation * scale; its planes matrix maps decal-space planes
		M4 dm = identity();
		dm.c[0] = extend(d.axis[0] * d.scale.x, 0.0f);
		dm.c[1] = extend(d.axis[1] * d.scale.y, 0.0f);
		dm.c[2] = extend(d.axis[2] * d.scale.z, 0.0f);
		dm.c[3] = extend(d.pos, 1.0f);
		M4 inv_t;
		if(!invert(transpose(dm), inv_t))
			inv_t = identity();
		ModelFrameData::DecalInfo info;
		info.camera_planes_matrix = mul(planes, inv_t);
		LightCubeAccum acc;
		const float min_square_distance = dot(d.scale, d.scale) * 0.25f;
		for(const sw_light& l : f.lights)
		{
			const V3 to_light = v3(l.position[0], l.position[1], l.position[2]) - d.pos;
			const float square_dist = std::max(dot(to_light, to_light), min_square_distance);
			const float inv_square_dist = 1.0f / square_dist;
			if(inv_square_dist < l.inv_square_radius)
				continue;
			const float k = inv_square_dist - l.inv_square_radius;
			const float light[3] = {l.color[0] * k, l.color[1] * k, l.color[2] * k};
			acc.add_sample(to_light, light);
		}
		info.dynamic_light = acc.convert();
		md.decal_info.push_back(info);

		V3 mn = d.pos, mx = d.pos;
		for(int k = 0; k < 8; ++k)
		{
			const V3 q = d.pos + d.axis[0] * ((k & 1) ? d.scale.x : -d.scale.x) + d.axis[1] * ((k & 2) ? d.scale.y : -d.scale.y) +
				d.axis[2] * ((k & 4) ? d.scale.z : -d.scale.z);
			mn = v3(std::min(mn.x, q.x), std::min(mn.y, q.y), std::min(mn.z, q.z));
			mx = v3(std::max(mx.x, q.x), std::max(mx.y, q.y), std::max(mx.z, q.z));
		}
		leafs_of_box(m, mn, mx, leafs);
		for(uint32_t l : leafs)
			md.leaf_decals[l].push_back(index);
	}
}

// ---------------------------------------------------------------------------------------------
// Order of a leaf's objects (draw_ordering.rs:86-226, LeafObjectsSorter)
// ---------------------------------------------------------------------------------------------
// A box seen from the camera: of each pair of opposite sides the one facing the camera (as a
// camera-space plane), and the eight corners in camera space.
struct ProjectedBBox
{
	Plane planes[3];
	V3 vertices[8];
};

Plane project_plane(const Plane& p, const M4& planes_matrix)
{
	const V4 t = mul(planes_matrix, extend(p.vec, -p.dist));
	return Plane{truncate(t), -t.w};
}

// project_bbox (draw_ordering.rs:74-80, 111-170)
ProjectedBBox project_bbox(const BBox& b, const M4& view, const M4& planes_matrix)
{
	ProjectedBBox r;
	const Plane pairs[3][2] = {{{v3(1, 0, 0), b.max.x}, {v3(-1, 0, 0), -b.min.x}}, {{v3(0, 1, 0), b.max.y}, {v3(0, -1, 0), -b.min.y}},
		{{v3(0, 0, 1), b.max.z}, {v3(0, 0, -1), -b.min.z}}};
	for(int i = 0; i < 3; ++i)
	{
		const Plane p0 = project_plane(pairs[i][0], planes_matrix), p1 = project_plane(pairs[i][1], planes_matrix);
		r.planes[i] = p0.dist < 0.0f ? p0 : p1;
	}
	// BBox::get_corners_vertices order (bbox.rs:83-95)
	for(int i = 0; i < 8; ++i)
		r.vertices[i] = view_transform(
			view, v3((i & 4) ? b.max.x : b.min.x, (i & 2) ? b.max.y : b.min.y, (i & 1) ? b.max.z : b.min.z));
	return r;
}

// is_bbox_at_front_of_plane (:198-215): a camera-facing plane with every corner at or in front of it
bool box_in_front_of_plane(const Plane& p, const V3 v[8])
{
	if(p.dist >= 0.0f)
		return false;
	for(int i = 0; i < 8; ++i)
		if(!(dot(p.vec, v[i]) >= p.dist))
			return false;
	return true;
}

float box_min_z(const ProjectedBBox& b)
{
	float r = b.vertices[0].z;
	for(int i = 0; i < 8; ++i)
		r = std::fmin(r, b.vertices[i].z);
	return r;
}

// compare_projected_bboxes (:172-196): true when `l` must be drawn after `r`
bool draw_after(const ProjectedBBox& l, const ProjectedBBox& r)
{
	for(const Plane& lp : l.planes)
		if(box_in_front_of_plane(lp, r.vertices))
			return false;
	for(const Plane& rp : r.planes)
		if(box_in_front_of_plane(rp, l.vertices))
			return true;
	return box_min_z(l) < box_min_z(r);
}

// order_bboxes (:82-109): the reference's pairwise exchange sort (no total order exists), at most
// 48 objects (MAX_OBJECTS, :48)
void order_objects(std::vector<std::pair<uint32_t, ProjectedBBox>>& objs)
{
	if(objs.size() > 48)
		objs.resize(48);
	for(size_t i = 0; i + 1 < objs.size(); ++i)
		for(size_t j = i + 1; j < objs.size(); ++j)
			if(draw_after(objs[i].second, objs[j].second))
				std::swap(objs[i], objs[j]);
}

// ---------------------------------------------------------------------------------------------
// Inline models in a leaf (partial_renderer.rs:3098-3288)
// ---------------------------------------------------------------------------------------------
// clip_3d_polygon_by_plane (clipping.rs:53-103): the part with dot(vec, v) >= dist
uint32_t clip_polygon_by_plane(const V3* poly, uint32_t n, const Plane& plane, V3* out, uint32_t cap)
{
	if(n == 0)
		return 0;
	const V3* prev = &poly[n - 1];
	float prev_d = dot(plane.vec, *prev);
	uint32_t count = 0;
	auto cut = [&](const V3& v0, const V3& v1) {
		// get_line_plane_intersection (clipping.rs:177-185)
		const float d0 = dot(v0, plane.vec) - plane.dist, d1 = dot(v1, plane.vec) - plane.dist, sum = d1 - d0;
		return v0 * (d1 / sum) - v1 * (d0 / sum);
	};
	for(uint32_t k = 0; k < n; ++k)
	{
		const V3& v = poly[k];
		const float d = dot(plane.vec, v);
		if(d > plane.dist)
		{
			if(prev_d < plane.dist)
			{
				out[count++] = cut(*prev, v);
				if(count == cap)
					break;
			}
			out[count++] = v;
			if(count == cap)
				break;
		}
		else if(d == plane.dist)
		{
			out[count++] = v;
			if(count == cap)
				break;
		}
		else if(prev_d > plane.dist)
		{
			out[count++] = cut(*prev, v);
			if(count == cap)
				break;
		}
		prev = &v;
		prev_d = d;
	}
	return count;
}

// draw_submodel_polygon (:3199-3288): the polygon's camera-space vertices clipped by the leaf's
// planes, skipping a plane the polygon lies on (and dropping the polygon when it lies on a plane
// facing the other way: the neighbour leaf draws it), then drawn with the leaf bounds' planes.
void draw_submodel_polygon(swh_frame& f, const swh_map& m, const V3 clip_planes[8], const Plane* leaf_planes,
	const Plane* leaf_planes_world, uint32_t num_leaf_planes, uint32_t pi, const int32_t clip_rect[4])
{
	const swh_frame::PolyData& pd = f.poly[pi];
	if(pd.visible_frame != f.frame_number)
		return;
	const Polygon& polygon = m.polygons[pi];
	const swh_frame::Basis& basis = f.basis[pi];
	V3 a[MAX_VERTICES], b[MAX_VERTICES];
	uint32_t n = std::min(polygon.num_vertices, MAX_VERTICES);
	for(uint32_t k = 0; k < n; ++k)
		a[k] = f.vertices_transformed[polygon.first_vertex + k];
	V3* src = a;
	V3* dst = b;
	const float almost_one = 1.0f - 1.0f / 256.0f, dist_eps = 1.0f / 4.0f;
	for(uint32_t i = 0; i < num_leaf_planes; ++i)
	{
		const Plane& w = leaf_planes_world[i];
		const float nd = dot(basis.normal, w.vec);
		if(nd >= almost_one)
		{
			if(std::fabs(dot(w.vec, basis.start) - w.dist) <= dist_eps)
				continue;
		}
		else if(nd <= -almost_one)
		{
			if(std::fabs(dot(w.vec, basis.start) - w.dist) <= dist_eps)
				return;
		}
		n = clip_polygon_by_plane(src, n, leaf_planes[i], dst, MAX_VERTICES);
		if(n < 3)
			return;
		std::swap(src, dst);
	}
	draw_polygon(f, clip_planes, src, n, pd.de, pd.te, pd.size, pd.surface_offset, SW_TEXSRC_SURFACE,
		m.materials[polygon.texture].blending_mode, clip_rect);
}

// draw_submodel_bsp_node_r (:3126-3197): the far child, the node's polygons (and the leaf's decals
// on them), the near child
void draw_submodel_node(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, const V4& planes_w, const V3 clip_planes[8],
	const Plane* leaf_planes, const Plane* leaf_planes_world, uint32_t num_leaf_planes, const std::vector<uint32_t>* leaf_decals,
	uint32_t node_index, const int32_t clip_rect[4], uint32_t depth)
{
	const SubmodelNode& node = m.submodel_nodes[node_index];
	const uint32_t num_nodes = uint32_t(m.submodel_nodes.size());
	const float w = dot(planes_w, extend(node.plane.vec, -node.plane.dist));
	const uint32_t mask = w >= 0.0f ? 1u : 0u;
	const uint32_t c_b = node.children[mask], c_f = node.children[mask ^ 1u];
	if(c_b < num_nodes && depth < 256)
		draw_submodel_node(f, m, cfg, planes_w, clip_planes, leaf_planes, leaf_planes_world, num_leaf_planes, leaf_decals, c_b,
			clip_rect, depth + 1);
	for(uint32_t pi = node.first_polygon; pi < node.first_polygon + node.num_polygons; ++pi)
	{
		draw_submodel_polygon(f, m, clip_planes, leaf_planes, leaf_planes_world, num_leaf_planes, pi, clip_rect);
		if(leaf_decals && !leaf_decals->empty())
			draw_polygon_decals(f, m, cfg, clip_planes, pi, *leaf_decals, clip_rect);
	}
	if(c_f < num_nodes && depth < 256)
		draw_submodel_node(f, m, cfg, planes_w, clip_planes, leaf_planes, leaf_planes_world, num_leaf_planes, leaf_decals, c_f,
			clip_rect, depth + 1);
}

// draw_view_models (partial_renderer.rs:2284-2309): view models over the viewport part, with no 3D
// clip planes but the near plane
void draw_view_models(swh_frame& f, const swh_map& m, const ClippingPolygon& viewport, const int32_t clip_rect[4])
{
	for(uint32_t vi : f.models.view_models)
		draw_mesh(f, m, viewport, nullptr, 0, f.models.visible[vi], clip_rect);
}

// draw_view_portal (partial_renderer.rs:3720-3776): the portal polygon in camera space, clipped by the
// leaf's planes, drawn with its portal texture (the view seen through it) when it is visible
void draw_view_portal(swh_frame& f, const swh_map& m, const V3 clip_planes[8], const Plane* leaf_planes, uint32_t num_leaf_planes,
	const M4& view, uint32_t pi, const int32_t clip_rect[4])
{
	const swh_frame::PortalInfo& info = f.portal_info[pi];
	if(!info.visible)
		return;
	const ViewPortalDef& portal = m.view_portals[pi];
	V3 a[MAX_VERTICES], b[MAX_VERTICES];
	uint32_t n = 4;
	for(uint32_t k = 0; k < n; ++k)
		a[k] = view_transform(view, portal.vertices[k]);
	V3* src = a;
	V3* dst = b;
	for(uint32_t i = 0; i < num_leaf_planes; ++i)
	{
		n = clip_polygon_by_plane(src, n, leaf_planes[i], dst, MAX_VERTICES);
		if(n < 3)
			return;
		std::swap(src, dst);
	}
	draw_polygon(f, clip_planes, src, n, info.de, info.te, info.resolution, info.texture_offset, SW_TEXSRC_PORTAL,
		portal.blending_mode, clip_rect);
}

void draw_leaf_objects(swh_frame& f, const swh_map& m, const swh_frame_config& cfg, uint32_t leaf, const ClippingPolygon& bounds,
	const int32_t clip_rect[4], const M4& view, const M4& planes)
{
	ModelFrameData& md = f.models;
	V3 clip_planes[8];
	bounds.clip_planes(clip_planes);
	const Leaf& lf = m.leafs[leaf];
	if(!md.leaf_decals.empty() && !md.leaf_decals[leaf].empty())
		for(uint32_t k = 0; k < lf.num_polygons; ++k)
			draw_polygon_decals(f, m, cfg, clip_planes, lf.first_polygon + k, md.leaf_decals[leaf], clip_rect);
	static const std::vector<uint32_t> no_objects;
	const std::vector<uint32_t>& submodels = f.leaf_submodels.empty() ? no_objects : f.leaf_submodels[leaf];
	const std::vector<uint32_t>& meshes = md.leaf_models.empty() ? no_objects : md.leaf_models[leaf];
	const std::vector<uint32_t>& portals = f.leaf_portals.empty() ? no_objects : f.leaf_portals[leaf];
	const size_t total = submodels.size() + meshes.size() + portals.size();
	if(total == 0)
		return;
	// the leaf's volume (a grid cell) as six inward planes, in world and camera space
	const BBox& b = m.leaf_bbox[leaf];
	const Plane world[6] = {{v3(1, 0, 0), b.min.x}, {v3(-1, 0, 0), -b.max.x}, {v3(0, 1, 0), b.min.y}, {v3(0, -1, 0), -b.max.y},
		{v3(0, 0, 1), b.min.z}, {v3(0, 0, -1), -b.max.z}};
	Plane cam[6];
	for(int i = 0; i < 6; ++i)
		cam[i] = project_plane(world[i], planes);
	const std::vector<uint32_t>* leaf_decals = md.leaf_decals.empty() ? nullptr : &md.leaf_decals[leaf];
	// the reference's object ids (:2663-2667); the front-end places no sprites
	const uint32_t SUBMODEL_ADD = 65536, MESH_ADD = SUBMODEL_ADD + 65536, PORTAL_ADD = MESH_ADD + 2 * 65536;
	auto draw_object = [&](uint32_t id) {
		if(id >= PORTAL_ADD)
			draw_view_portal(f, m, clip_planes, cam, 6, view, id - PORTAL_ADD, clip_rect);
		else if(id >= MESH_ADD)
			draw_mesh(f, m, bounds, cam, 6, md.visible[id - MESH_ADD], clip_rect);
		else
		{
			const uint32_t k = id - SUBMODEL_ADD;
			const swh_frame::SubmodelFrame& s = f.submodels[k];
			if(s.active)
				draw_submodel_node(f, m, cfg, s.planes.row(3), clip_planes, cam, world, 6, leaf_decals, m.submodels[k].root_node,
					clip_rect, 0);
		}
	};
	if(total == 1)
	{
		draw_object(!submodels.empty() ? SUBMODEL_ADD + submodels[0] : !meshes.empty() ? MESH_ADD + meshes[0] : PORTAL_ADD + portals[0]);
		return;
	}
	// several objects: the reference's projected-box order (draw_leaf, :2646-2714)
	std::vector<std::pair<uint32_t, ProjectedBBox>> objs;
	for(uint32_t k : submodels)
	{
		const swh_frame::SubmodelFrame& s = f.submodels[k];
		if(!s.active)
			continue;
		// get_model_bbox_for_ordering (inline_models_index.rs:66-80): shrunk by 1 towards the centre
		BBox ob = s.bbox;
		const V3 c = (ob.min + ob.max) * 0.5f;
		ob.max = v3(std::fmax(ob.max.x - 1.0f, c.x), std::fmax(ob.max.y - 1.0f, c.y), std::fmax(ob.max.z - 1.0f, c.z));
		ob.min = v3(std::fmin(ob.min.x + 1.0f, c.x), std::fmin(ob.min.y + 1.0f, c.y), std::fmin(ob.min.z + 1.0f, c.z));
		objs.push_back({SUBMODEL_ADD + k, project_bbox(ob, s.view, s.planes)});
	}
	for(uint32_t vi : meshes)
		objs.push_back({MESH_ADD + vi, project_bbox(md.visible[vi].bbox, md.visible[vi].view, md.visible[vi].planes)});
	for(uint32_t pi : portals)
	{
		// every portal placed in the leaf takes part in the ordering, visible or not (:2724-2737)
		const ViewPortalDef& portal = m.view_portals[pi];
		BBox pb{portal.vertices[0], portal.vertices[0]};
		for(int k = 1; k < 4; ++k)
		{
			const V3 v = portal.vertices[k];
			pb.min = v3(std::fmin(pb.min.x, v.x), std::fmin(pb.min.y, v.y), std::fmin(pb.min.z, v.z));
			pb.max = v3(std::fmax(pb.max.x, v.x), std::fmax(pb.max.y, v.y), std::fmax(pb.max.z, v.z));
		}
		objs.push_back({PORTAL_ADD + pi, project_bbox(pb, view, planes)});
	}
	order_objects(objs);
	for(const auto& o : objs)
		draw_object(o.first);
}

} // namespace swh
